// Fused float32 pipeline for Gaussian and Beta primitives (the working-size path).
//
// Forward:  fused projection (parameter transform + projection + double
//           precision conic/mu + radii + tile count, one 32-byte splat record
//           per (camera, primitive)) -> device scan -> binning into packed
//           camera-major tile lists -> rasterization. Binning places each
//           (record, tile) pair directly (per-range bin counts, a scan, then
//           placement) when a camera's (tile, height class) bins fit in a
//           small LDS table, and otherwise emits (tile, record) pairs, sorts
//           them with a capacity-bounded stable radix sort and packs the
//           records per list.
// Backward: rasterization VJP, one partial per intersection stored at the
//           intersection's emission index (the emission order groups each
//           (camera, primitive)'s intersections contiguously) -> one lane per
//           primitive summing its partials and the conic/mu, projection and
//           parameter VJPs over cameras, all in a fixed order: the backward
//           uses no atomics and is bitwise reproducible.
// All sizes that depend on the intersection count are read on the device, so
// the whole sequence can be captured into a hipGraph; the only host-visible
// condition is capacity overflow (reported in a device flag).
//
// Semantics are those of the operator kernels (raster.hip, geometry.hip) with
// two precomputations per primitive: sigma is evaluated in log2 units
// (A = a/2 log2 e, B = b log2 e, C = c/2 log2 e, q = sigma log2 e) and the
// contribution test "sigma >= 0 and |w e^-sigma| >= 1e-5" becomes
// 0 <= q <= S with S = log2(|w| 1e5). Non-negative floats order like their
// bit patterns and every negative float or NaN has a larger one, so the test
// is the single unsigned compare bits(q) < qlim with qlim = bits(S) + 1, and
// qlim = 0 (nothing passes) when S < 0, w = 0 or S is NaN.
// Beta primitives (d^2 = delta^T P delta, alpha = w (1 - d^2)^b2d kept for
// 0 <= d^2 < 1 and |alpha| >= 1e-5) use the same single compare on d^2 with
// the bound D = min(largest float below 1, 1 - (1e-5 / |w|)^(1 / b2d))
// (computed in double); the backward decides the pairs in a narrow band
// around D, where float32 d^2 may fall on the other side from float64 d^2, in
// double from the record's double-precision form (BetaEdge, edge_keep).
#include <algorithm>
#include <type_traits>

#include "binning.hpp"
#include "common.hpp"
#include "geometry_common.hpp"
#include "raster_common.hpp"

namespace r2s {

// Gaussian record: sigma log2 e = A dx^2 + B dx dy + C dy^2, alpha = w 2^-q.
struct __align__(16) Splat {
    float mx, my, A, B, C, w;
    uint32_t qlim;  // bits(S) + 1, or 0 when the primitive contributes nowhere
    int32_t g;      // first emission index (projected) / emission index (packed)
};
static_assert(sizeof(Splat) == 32, "splat record must be 32 bytes");

// Wide record, for Beta primitives and for primitives with an uncertainty:
// Beta: d^2 = A dx^2 + B dx dy + C dy^2 (A = a, B = 2 b, C = c of the conic),
// alpha = w (1 - d^2)^b2d with w = opacity mu ratio(beta), b2d = beta + 1/2,
// and the mean is mx + lx, my + ly (its float32 value and remainder, from the
// projection in double; offsets are formed as (p - mx) - lx); Gaussian with
// uncertainty: the Splat fields (b2d, lx, ly unused). unc is the primitive's
// uncertainty q (variance = sum q alpha^2).
struct __align__(16) BetaSplat {
    float mx, my, A, B, C, w;
    uint32_t qlim;  // Beta: bits(D) + 1; Gaussian: bits(S) + 1; or 0
    int32_t g;
    float b2d, unc, lx, ly;
};
static_assert(sizeof(BetaSplat) == 48, "wide record must be 48 bytes");

template <bool BETA, bool VAR = false>
using Rec = typename std::conditional<BETA || VAR, BetaSplat, Splat>::type;

// Double-precision form of a Beta record: its mean, conic (d^2 = a dx^2 +
// 2 b dx dy + c dy^2) and the largest d^2 whose term passes the cut-off
// (|alpha| = |w| (1 - d^2)^b2d >= 1e-5; -1 when none does). The raster
// backward reads it only for pairs whose float32 d^2 lies within rounding of
// the record's float32 bound (edge_keep), so that each of its kept / skipped
// decisions is the one float64 makes: at the cut-off a Beta term's gradient
// jumps by about w b2d h^(b2d - 1) |d d^2|, not by 1e-5 as a Gaussian's does.
// (The forward keeps the float32 decision: a pair decided the other way moves
// the image by at most 1e-5.)
struct __align__(16) BetaEdge {
    double mx, my, a, b, c, dmax;
};
// Stored Beta record: the 48-byte record the raster kernels gather, then its
// double-precision form.
struct __align__(16) BetaRec {
    BetaSplat s;
    BetaEdge e;
};
static_assert(sizeof(BetaRec) == 96, "stored Beta record must be 96 bytes");

template <bool BETA, bool VAR = false>
using Stored = typename std::conditional<BETA, BetaRec, Rec<BETA, VAR>>::type;

// The packed buffer the forward leaves for the raster kernels and the
// backward: `capacity` list entries {record index, emission index} (uint2),
// then (from a 256-byte boundary) the C*N projected records.
__host__ __device__ constexpr size_t entries_bytes(uint32_t capacity) {
    return (size_t(capacity) * sizeof(uint2) + 255) / 256 * 256;
}

template <typename R>
__device__ __forceinline__ const R& head(const R& r) { return r; }
__device__ __forceinline__ const BetaSplat& head(const BetaRec& r) { return r.s; }

// Pixel-minus-mean offsets of a record: Beta records carry their mean's
// float32 remainder (the profile's slope at its support edge would pass the
// rounding of a pixel coordinate to the image), the others do not.
template <bool BETA, typename R>
__device__ __forceinline__ float offset_x(const R& s, float px) {
    if constexpr (BETA)
        return (px - s.mx) - s.lx;
    else
        return px - s.mx;
}
template <bool BETA, typename R>
__device__ __forceinline__ float offset_y(const R& s, float py) {
    if constexpr (BETA)
        return (py - s.my) - s.ly;
    else
        return py - s.my;
}

constexpr float kLog2e = 1.4426950408889634f;
constexpr float kLog2Inv1e5 = 16.609640474436812f;  // log2(1e5)
constexpr int kDefaultWaves = 8;                     // waves sharing one tile unless chosen otherwise
// Floats per intersection partial, a 32-byte slot: (v_mx, v_my, v_a, v_b,
// v_c, v_w, v_b2d, v_q), the last two 0 where the kind has no Beta exponent or
// uncertainty. (Packing plain Gaussians' six floats into 24 bytes measured
// ~80 us slower in the parameter VJP at 27M intersections: 8-byte loads and
// LDS reads in place of 16-byte ones.)
constexpr int kPartialStride = 8;

__device__ __forceinline__ uint32_t qlim_of(float w) {
    const float S = __log2f(fabsf(w)) + kLog2Inv1e5;
    return (w != 0.f && S >= 0.f) ? __float_as_uint(S) + 1u : 0u;
}

__device__ __forceinline__ bool contributes(float q, uint32_t qlim) { return __float_as_uint(q) < qlim; }

// The Beta line-integral factor r(beta) = Gamma(beta + 1) / (sqrt(2)
// Gamma(beta + 3/2)) in double, as the host computes it (lgamma below 20; from
// 20 the Stirling series of log Gamma(z) - log Gamma(z + 1/2) + 1/2 log z, z =
// beta + 1, which avoids the cancellation of two large lgamma values).
__device__ __forceinline__ double beta_ratio_f64(double beta) {
    constexpr double kHalfLn2 = 0.34657359027997264;
    if (beta < 20.0) return exp(lgamma(beta + 1.0) - lgamma(beta + 1.5) - kHalfLn2);
    const double z = beta + 1.0, w = 1.0 / z, w2 = w * w;
    const double series = w * (1.0 / 8.0 + w2 * (-1.0 / 192.0 + w2 * (1.0 / 640.0 + w2 * (-17.0 / 14336.0 +
                                                                                          w2 * (31.0 / 18432.0)))));
    return exp(series - 0.5 * (2.0 * kHalfLn2 + log(z)));
}

// The Beta cut-off in double: the largest d^2 with |w| (1 - d^2)^b2d >= 1e-5
// (-1 when |w| < 1e-5), and the qlim of its float32 rounding.
// Takes inv_b2d = 1 / b2d (NaN when b2d <= 0) and forms the bound as
// 1 - 2^((log2 1e-5 - log2 |w|) inv_b2d).
__device__ __forceinline__ double beta_dmax_inv(double w, double inv_b2d) {
    constexpr double kLog2Thr = -16.609640474436812;  // log2(1e-5)
    const double aw = fabs(w);
    if (!(aw >= 1e-5) || !(inv_b2d > 0.0)) return -1.0;
    return 1.0 - exp2((kLog2Thr - log2(aw)) * inv_b2d);
}
__device__ __forceinline__ uint32_t beta_qlim_from(double dmax) {
    if (!(dmax >= 0.0)) return 0u;
    return __float_as_uint(fminf(float(dmax), 0.99999994f)) + 1u;
}

// Half-width of the band around a Beta record's float32 bound D in which its
// float32 d^2 can fall on the other side of the cut-off from the float64
// d^2: with rho = |b| / sqrt(a c) the terms of d^2 sum to at most (1 + rho) /
// (1 - rho) d^2, each carrying a few roundings (conic, offsets, products);
// 16 float32 ulps of that sum, plus the rounding of D itself. Degenerate
// forms (rho near 1) get the whole range.
__device__ __forceinline__ float edge_band(const BetaSplat& s) {
    const float ac = s.A * s.C, b = 0.5f * s.B;
    const float rho = (ac > 0.f) ? fabsf(b) * __builtin_amdgcn_rsqf(ac) : 1.f;
    if (!(rho < 0.999f)) return 2.f;
    return 9.5367431640625e-07f * (1.f + rho) / (1.f - rho) + 5.9604644775390625e-08f;  // 16 ulp(1) x ..., + ulp(1)
}

// The edge band as float32 bit limits: pairs with bits(d^2) < x pass the
// cut-off whatever their float64 d^2, pairs with bits(d^2) >= y fail it, and
// those in between are settled in double (edge_keep). Both 0 when the record
// contributes nowhere (and for records without a Beta form).
template <typename R>
__device__ __forceinline__ uint2 edge_limits(const R& s) {
    if constexpr (std::is_same<R, BetaSplat>::value) {
        if (s.qlim == 0u) return make_uint2(0u, 0u);
        const float bound = __uint_as_float(s.qlim - 1u), band = edge_band(s);
        const float lo = bound - band, hi = bound + band;
        return make_uint2(lo >= 0.f ? __float_as_uint(lo) + 1u : 0u, __float_as_uint(hi) + 1u);
    } else {
        return make_uint2(0u, 0u);
    }
}

// Whether a Beta pair passes the cut-off in double: pixel centre (px, py),
// the record's double-precision form e (0 <= d^2 < 1 and d^2 <= dmax).
__device__ __forceinline__ bool edge_keep(const BetaEdge& e, float px, float py) {
    const double dx = e.mx - double(px), dy = e.my - double(py);
    const double q = e.a * dx * dx + 2.0 * e.b * dx * dy + e.c * dy * dy;
    return q >= 0.0 && q < 1.0 && q <= e.dmax;
}

// LDS written by some lanes of a wave is visible to its other lanes.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The record's quadratic form q (Gaussian: sigma log2 e; Beta: d^2) at one
// pixel, in the two evaluation orders of the raster kernels. The forward
// shares the column terms of a lane's pixels (adx2 = A dx dx, bdx = B dx with
// dx = px - mx, dy = py - my); the backward shares the row terms of a row's
// pixels (bdy = B dy, cdy2 = C dy dy with dy = (y0 + r) - my and dx =
// (x0 - mx) + j). Which pixels pass the cut-off is decided by these exact
// operations (fused_keep_mask_kernel reproduces both).
__device__ __forceinline__ float form_by_columns(float adx2, float bdx, float C, float dy) {
    return fmaf(dy, fmaf(C, dy, bdx), adx2);
}
__device__ __forceinline__ float form_by_rows(float A, float dx, float bdy, float cdy2) {
    return fmaf(dx, fmaf(A, dx, bdy), cdy2);
}

// Live intersection count; an overflowed forward (count > capacity) has an
// incomplete record list, so every consumer treats it as empty.
__device__ __forceinline__ uint32_t live(const uint32_t* __restrict__ d_n, uint32_t capacity) {
    const uint32_t n = *d_n;
    return n <= capacity ? n : 0u;
}

// Rows [lo, hi) of a T-row tile (pixel centres y_base + r + 1/2) that a record
// can reach. For every dx, q = A dx^2 + B dx dy + C dy^2 >= dy^2 (4AC - B^2) /
// (4A), so a row with dy^2 > 4 A S / (4AC - B^2) (S the record's bound: the
// Gaussian's S or the Beta's D, stored as qlim - 1) holds no contributing pixel.
// The extent is widened (1e-3 relative, 0.05 px) to cover the rounding of q;
// forms that are not positive definite get the whole tile, records that
// contribute nowhere (qlim = 0) none.
template <typename R>
__device__ __forceinline__ void row_span(const R& s, float y_base, int T, int& lo, int& hi) {
    lo = 0;
    hi = T;
    if (s.qlim == 0u) {
        hi = 0;
        return;
    }
    const float S = __uint_as_float(s.qlim - 1u);
    const float den = 4.f * s.A * s.C - s.B * s.B;
    // hardware reciprocal and square root (about 1 ulp): covered by the margin
    const float e2 = 4.f * s.A * S * __builtin_amdgcn_rcpf(den);
    if (!(s.A > 0.f) || !(den > 0.f) || !(e2 < 1e30f)) return;
    const float ext = __builtin_amdgcn_sqrtf(e2) * 1.001f + 0.05f;
    const float c = s.my - y_base - 0.5f;  // the mean in row units of this tile
    const float fl = fminf(fmaxf(ceilf(c - ext), 0.f), float(T));
    const float fh = fminf(fmaxf(floorf(c + ext) + 1.f, 0.f), float(T));
    lo = int(fl);
    hi = max(lo, int(fh));
}

// Tile rectangle and row reach of one projected record. The projection
// writes the rectangle (uint2: x0 | x1 << 16, y0 | y1 << 16, tiles [x0, x1) x
// [y0, y1), 0 when culled) and, for the direct-placement binning, the reach
// (float2: the mean row and row_span's half-extent, -1 reaching no row and
// +inf every row of a tile), so the binning never reads a record to classify
// an intersection by height.
struct SpanRect {
    uint32_t xr, yr;
    float my, ext;
};

__device__ __forceinline__ SpanRect load_span(const uint2* __restrict__ rects, const float2* __restrict__ reach,
                                              size_t i) {
    const uint2 r = rects[i];
    const float2 e = reach[i];
    return SpanRect{r.x, r.y, e.x, e.y};
}

// row_span's extent of a record as a value (-1: no row; +inf: whole tile).
template <typename R>
__device__ __forceinline__ float row_reach(const R& s) {
    if (s.qlim == 0u) return -1.f;
    const float S = __uint_as_float(s.qlim - 1u);
    const float den = 4.f * s.A * s.C - s.B * s.B;
    const float e2 = 4.f * s.A * S * __builtin_amdgcn_rcpf(den);
    if (!(s.A > 0.f) || !(den > 0.f) || !(e2 < 1e30f)) return __builtin_inff();
    return __builtin_amdgcn_sqrtf(e2) * 1.001f + 0.05f;
}

// Number of tile rows [lo, hi) of a record with mean row my and extent ext
// that a T-row tile starting at y_base holds (row_span's result).
__device__ __forceinline__ uint32_t rows_reached(float my, float ext, float y_base, uint32_t T) {
    if (ext < 0.f) return 0u;
    const float c = my - y_base - 0.5f;
    const float fl = fminf(fmaxf(ceilf(c - ext), 0.f), float(T));
    const float fh = fminf(fmaxf(floorf(c + ext) + 1.f, 0.f), float(T));
    return fh > fl ? uint32_t(fh - fl) : 0u;
}

// ---------------------------------------------------------------------------
// Primitive order. The forward processes primitives in the order of a 15-bit
// Morton code of their means (5 bits per axis over the means' bounding box,
// 32^3 cells; ties keep the input order; two radix passes), or of its top 8
// bits for sets of up to 2^17 primitives (one pass): consecutive primitives then project to nearby
// tiles in every view, so the binning's gathers and writes stay local. The
// records, the per-record intersection offsets and the emission order all
// follow that order; the parameter VJP maps back to the input order.
// ---------------------------------------------------------------------------
constexpr uint32_t kOrderThreads = 1024;

// Floats as unsigned integers of the same order (for min/max atomics).
__device__ __forceinline__ uint32_t order_bits(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float order_float(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// bounds (ordered bits) = {min x, min y, min z, max x, max y, max z} of the
// finite means; bounds starts at {~0 x3, 0 x3} (morton_bounds_init_kernel).
__global__ void morton_bounds_init_kernel(uint32_t* __restrict__ bounds) {
    if (threadIdx.x < 6) bounds[threadIdx.x] = threadIdx.x < 3 ? 0xffffffffu : 0u;
}

__global__ void __launch_bounds__(kOrderThreads)
morton_bounds_kernel(const float* __restrict__ means, uint32_t N, uint32_t* __restrict__ bounds) {
    __shared__ float red[6][kOrderThreads / kWave];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t n = blockIdx.x * kOrderThreads + threadIdx.x; n < N; n += gridDim.x * kOrderThreads) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float x = means[3 * size_t(n) + k];
            if (isfinite(x)) {
                lo[k] = fminf(lo[k], x);
                hi[k] = fmaxf(hi[k], x);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        for (uint32_t off = kWave / 2; off > 0; off >>= 1) {
            lo[k] = fminf(lo[k], __shfl_xor(lo[k], off, kWave));
            hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], off, kWave));
        }
    }
    const uint32_t wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            red[k][wave] = lo[k];
            red[3 + k][wave] = hi[k];
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const bool is_lo = threadIdx.x < 3;
        float v = red[threadIdx.x][0];
        for (uint32_t w = 1; w < kOrderThreads / kWave; ++w)
            v = is_lo ? fminf(v, red[threadIdx.x][w]) : fmaxf(v, red[threadIdx.x][w]);
        if (is_lo) atomicMin(bounds + threadIdx.x, order_bits(v));
        else atomicMax(bounds + threadIdx.x, order_bits(v));
    }
}

constexpr uint32_t kMortonAxisBits = 5;
constexpr uint32_t kMortonCoarseMax = 1u << 17;  // primitive sets sorted by the top 8 code bits

__device__ __forceinline__ uint32_t spread_bits3(uint32_t v) {  // up to 10 bits -> every third bit
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void __launch_bounds__(256)
morton_keys_kernel(const float* __restrict__ means, uint32_t N, const uint32_t* __restrict__ bounds,
                   uint32_t drop, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    uint32_t key = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        // no finite mean: lo = NaN and every key 0
        const float lo = order_float(bounds[k]), ext = order_float(bounds[3 + k]) - lo;
        const float t = ext > 0.f ? (means[3 * size_t(n) + k] - lo) / ext : 0.f;
        constexpr float kCells = float(1u << kMortonAxisBits);
        const float q = fminf(fmaxf(t * kCells, 0.f), kCells - 1.f);  // NaN -> 0
        key |= spread_bits3(uint32_t(q)) << (2 - k);
    }
    keys[n] = key >> drop;
    vals[n] = n;
}

// The camera-independent double-precision terms of one primitive's
// projection: its 3D covariance (00, 01, 02, 11, 12, 22) from the normalised
// quaternion and the scales, as the float64 parameter transform computes and
// stores it, and (Beta) the line-integral ratio B(1/2, beta + 1) / sqrt(2 pi)
// and 1 / b2d (b2d = beta + 1/2; NaN when b2d <= 0) in double. 64 bytes.
struct PrimFrame {
    double cov[6];
    double ratio;
    double inv_b2d;
};

// One camera in double (exact from its float32 viewmat and K) with its
// field-of-view limits, computed once per camera for the projection.
struct CamFrame {
    Camera<double> cam;
    FovLimits<double> fov;
};

__device__ __forceinline__ void camera_frame(const float* __restrict__ viewmats, const float* __restrict__ Ks,
                                             uint32_t c, uint32_t W, uint32_t H, CamFrame* __restrict__ out) {
    CamFrame f;
    f.cam.load(viewmats + 16 * size_t(c), Ks + 9 * size_t(c));
    f.fov = fov_limits(f.cam, double(W), double(H));
    out[c] = f;
}

// Copies of the primitive parameters in the sorted order (perm[p] = input
// index), and each primitive's PrimFrame; workgroup 0 also writes the C
// camera frames.
__global__ void __launch_bounds__(256)
permute_params_kernel(const uint32_t* __restrict__ perm, uint32_t N, const float* __restrict__ means,
                      const float* __restrict__ quats, const float* __restrict__ scales,
                      const float* __restrict__ dens, const float* __restrict__ betas,
                      const float* __restrict__ ratio, const float* __restrict__ unc, float* __restrict__ o_means,
                      float* __restrict__ o_dens, float* __restrict__ o_betas, float* __restrict__ o_ratio,
                      float* __restrict__ o_unc, PrimFrame* __restrict__ o_frames,
                      const float* __restrict__ viewmats, const float* __restrict__ Ks, uint32_t C, uint32_t W,
                      uint32_t H, CamFrame* __restrict__ cams) {
    if (blockIdx.x == 0)
        for (uint32_t c = threadIdx.x; c < C; c += blockDim.x) camera_frame(viewmats, Ks, c, W, H, cams);
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const size_t n = perm[p];
#pragma unroll
    for (int k = 0; k < 3; ++k) o_means[3 * size_t(p) + k] = means[3 * n + k];
    o_dens[p] = dens[n];
    if (betas) {
        o_betas[p] = betas[n];
        o_ratio[p] = ratio[n];
    }
    if (unc) o_unc[p] = unc[n];
    const double raw[4] = {double(quats[4 * n]), double(quats[4 * n + 1]), double(quats[4 * n + 2]),
                           double(quats[4 * n + 3])};
    double q[4], mag, norm, R[9];
    unit_quat(raw, q, mag, norm);
    quat_to_rot(q, R);
    const double s[3] = {double(scales[3 * n]), double(scales[3 * n + 1]), double(scales[3 * n + 2])};
    PrimFrame f;
    covariance_from(R, s, f.cov);
    const double b2d = betas ? double(betas[n]) + 0.5 : 0.0;
    f.ratio = betas ? beta_ratio_f64(double(betas[n])) : 0.0;
    f.inv_b2d = b2d > 0.0 ? 1.0 / b2d : __builtin_nan("");
    o_frames[p] = f;
}

// Projected mean and 2D covariance of one primitive for one camera in double
// precision (its float32 mean is exact in double; cov is its PrimFrame
// covariance): the operator chain parameter transform -> world-to-image
// projection evaluated in double. Returns false when the primitive is culled
// (depth <= 0.01).
__device__ __forceinline__ bool project_f64(const CamFrame& cf, const float* __restrict__ mean, const double cov[6],
                                         double m2[2], double c2[6]) {
    const Camera<double>& cam = cf.cam;
    const double p[3] = {double(mean[0]), double(mean[1]), double(mean[2])};
    ProjState<double> st;
    if (!project_state(cam, cf.fov, p, cov, st)) return false;
    m2[0] = cam.fx * st.x / st.z + cam.cx;
    m2[1] = cam.fy * st.y / st.z + cam.cy;
    projected_cov(st, c2);
    return true;
}

// ---------------------------------------------------------------------------
// Fused projection + tile count. grid (ceil(N/256), ceil(C/kProjectCams)):
// each thread takes one primitive (its PrimFrame and parameters loaded once)
// for kProjectCams consecutive cameras. Each (camera, primitive) is projected
// in double (project_f64: culling, radii, tile rectangle, conic and mu as the
// operators compute them in float64) and its record stores the results
// rounded to float32; Beta records also keep their mean's float32 remainder
// and the cut-off in double.
// ---------------------------------------------------------------------------
constexpr uint32_t kProjectCams = 4;

template <bool BETA, bool VAR>
__global__ void __launch_bounds__(256)
fused_project_kernel(const float* __restrict__ means, const PrimFrame* __restrict__ frames,
                     const float* __restrict__ dens, const float* __restrict__ betas,
                     const float* __restrict__ ratio, const float* __restrict__ unc,
                     const CamFrame* __restrict__ cams, uint32_t C, uint32_t W,
                     uint32_t H, float eps, float clip, uint32_t tile_size, uint32_t ntx, uint32_t nty, uint32_t N,
                     Stored<BETA, VAR>* __restrict__ recs, uint2* __restrict__ rects, float2* __restrict__ reach,
                     uint32_t* __restrict__ counts) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const PrimFrame f = frames[n];
    const float density = dens[n];
    const float beta = BETA ? betas[n] : 0.f, beta_ratio = BETA ? ratio[n] : 0.f;
    const float uncertainty = VAR ? unc[n] : 0.f;
    const uint32_t c_end = min(C, (blockIdx.y + 1u) * kProjectCams);
#pragma unroll 1
    for (uint32_t c = blockIdx.y * kProjectCams; c < c_end; ++c) {
        const size_t idx = size_t(c) * N + n;
        Rec<BETA, VAR> rec = {};
        BetaEdge edge = {0.0, 0.0, 0.0, 0.0, 0.0, -1.0};
        rec.g = int32_t(idx);
        if constexpr (VAR) rec.unc = uncertainty;
        uint2 rect = make_uint2(0u, 0u);
        float2 rr = make_float2(0.f, -1.f);
        uint32_t count = 0;
        double m2[2], c2[6];
        if (project_f64(cams[c], means + 3 * size_t(n), f.cov, m2, c2)) {
            int32_t rx, ry;
            footprint_radii(c2, m2[0], m2[1], double(eps), double(clip), double(W), double(H), rx, ry);
            double P[3], mu;
            conic_mu_fast_f64(c2, double(eps), P, mu);
            rec.mx = float(m2[0]);
            rec.my = float(m2[1]);
            if constexpr (BETA) {
                rec.lx = float(m2[0] - double(rec.mx));
                rec.ly = float(m2[1] - double(rec.my));
                const float mus = float(mu) * beta_ratio;  // mu_beta = mu ratio(beta), in float32
                const float w = density * mus;
                rec.A = float(P[0]);
                rec.B = 2.f * float(P[1]);
                rec.C = float(P[2]);
                rec.w = w;
                rec.b2d = beta + 0.5f;
                // the cut-off in double (w = density mu r(beta) and b2d as float64 forms them)
                const double dmax = beta_dmax_inv(double(density) * mu * f.ratio, f.inv_b2d);
                rec.qlim = beta_qlim_from(dmax);
                edge = BetaEdge{m2[0], m2[1], P[0], P[1], P[2], dmax};
            } else {
                const float w = density * float(mu);
                rec.A = 0.5f * float(P[0]) * kLog2e;
                rec.B = float(P[1]) * kLog2e;
                rec.C = 0.5f * float(P[2]) * kLog2e;
                rec.w = w;
                rec.qlim = qlim_of(w);
            }
            uint32_t x0, x1, y0, y1;
            if (rx > 0 && ry > 0 && tile_span(m2[0], rx, tile_size, ntx, x0, x1) &&
                tile_span(m2[1], ry, tile_size, nty, y0, y1)) {
                count = (x1 - x0) * (y1 - y0);
                rect = make_uint2(x0 | (x1 << 16), y0 | (y1 << 16));
                rr = make_float2(rec.my, row_reach(rec));
            }
        }
        if constexpr (BETA)
            recs[idx] = BetaRec{rec, edge};
        else
            recs[idx] = rec;
        rects[idx] = rect;
        if (reach) reach[idx] = rr;
        counts[idx] = count;
    }
}

// Records (key = tile within the camera, ty * ntx + tx; value = cam * N + n)
// at their scanned positions, which are camera-major; entries that would pass
// the capacity raise the overflow flag. A stable sort on the key alone then
// orders the list by (tile, camera, primitive): one radix pass for up to 256
// tiles per camera.
__global__ void __launch_bounds__(256)
fused_emit_kernel(const uint2* __restrict__ rects, const uint32_t* __restrict__ cum, uint32_t total,
                  uint32_t ntx, uint32_t capacity, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                  uint32_t* __restrict__ overflow) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const uint32_t w0 = cum[i], cnt = cum[i + 1] - w0;
    if (cnt == 0) return;
    if (w0 + cnt > capacity) {
        *overflow = 1u;
        return;
    }
    const uint2 r = rects[i];
    const uint32_t x0 = r.x & 0xffffu, x1 = r.x >> 16, y0 = r.y & 0xffffu, y1 = r.y >> 16;
    uint32_t w = w0;
    for (uint32_t y = y0; y < y1; ++y)
        for (uint32_t x = x0; x < x1; ++x, ++w) {
            keys[w] = y * ntx + x;
            vals[w] = i;
        }
}

// Starts of the sorted (tile-major) lists: offsets[k] for list k = tile * C +
// cam is the first sorted intersection whose (tile, camera) is not below it
// (lower_bound on tile * C + value / N).
__global__ void __launch_bounds__(256)
fused_offsets_kernel(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                     const uint32_t* __restrict__ d_n, uint32_t capacity, uint32_t C, uint32_t N,
                     uint32_t n_lists, int32_t* __restrict__ offsets) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_lists) return;
    const uint32_t n = live(d_n, capacity);
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        const uint64_t probe = uint64_t(keys[mid]) * C + vals[mid] / N;
        if (probe < k) lo = mid + 1; else hi = mid;
    }
    offsets[k] = int32_t(lo);
}

// Tile (cam * nt + t) -> [start, end) of its packed list (camera-major
// offsets, the reference's layout). Wave-uniform.
__device__ __forceinline__ void fused_tile(const int32_t* __restrict__ offsets, uint32_t tile, uint32_t n_tiles,
                                           uint32_t n_live, int32_t& start, int32_t& end) {
    start = offsets[tile];
    end = (tile + 1 == n_tiles) ? int32_t(n_live) : offsets[tile + 1];
    start = max(0, min(start, int32_t(n_live)));
    end = max(start, min(end, int32_t(n_live)));
}

// The raster kernels deal their n tiles to the 8 XCDs (workgroup b runs on
// XCD b % 8) as contiguous bands of the camera-major tile index: band x holds
// per + (x < rem) tiles (per = n / 8, rem = n % 8), so the tiles that gather
// the same records (neighbours in one view) share one XCD's L2. xcd_slot is
// the position of workgroup b in the band-major order.
__device__ __forceinline__ uint32_t xcd_slot(uint32_t b, uint32_t n) {
    const uint32_t per = n / 8u, rem = n % 8u, x = b % 8u;
    return x * per + min(x, rem) + b / 8u;
}

// The tile order: each XCD band's tiles (the band the raster launches give
// XCD x: per + (x < rem) tiles of the camera-major index from x per +
// min(x, rem)) in decreasing size class of 2^(1/8) (class 255 - floor(8
// log2(count + 1))), so that each XCD walks its tiles longest first and the
// few heavy tiles of a view do not all land on the same compute units at the
// end of the launch. One workgroup per band deals its tiles into the classes
// with an LDS histogram, one wave's exclusive scan of it and atomic slots.
// The order inside a class is whatever the atomics give: it decides only
// which workgroup runs a tile when, never a result.
constexpr uint32_t kTileClasses = 256, kTileOrderThreads = 1024;

// Tile t's list starts at starts[t * stride] (stride 1: the tile offsets;
// kBinClasses: the direct placement's scanned bin starts, a tile's first bin)
// and ends where tile t + 1's starts (the last at the live count).
__device__ __forceinline__ uint32_t tile_class(const uint32_t* __restrict__ starts, uint32_t stride, uint32_t t,
                                               uint32_t n_tiles, uint32_t n_live) {
    const uint32_t start = min(starts[size_t(t) * stride], n_live);
    const uint32_t end = max(start, min(t + 1 == n_tiles ? n_live : starts[size_t(t + 1) * stride], n_live));
    const float cls = floorf(8.f * __log2f(float(end - start) + 1.f));
    return 255u - min(255u, uint32_t(cls));
}

// offsets_out (when given) also receives the tile offsets, starts[t * stride].
__global__ void __launch_bounds__(kTileOrderThreads)
fused_tile_order_kernel(const uint32_t* __restrict__ starts, uint32_t stride, const uint32_t* __restrict__ d_n,
                        uint32_t capacity, uint32_t n_tiles, uint32_t* __restrict__ order,
                        int32_t* __restrict__ offsets_out) {
    __shared__ uint32_t slot[kTileClasses];
    const uint32_t x = blockIdx.x, per = n_tiles / 8u, rem = n_tiles % 8u;
    const uint32_t lo = x * per + min(x, rem), hi = lo + per + (x < rem ? 1u : 0u);
    const uint32_t n_live = live(d_n, capacity);
    for (uint32_t c = threadIdx.x; c < kTileClasses; c += blockDim.x) slot[c] = 0u;
    __syncthreads();
    for (uint32_t t = lo + threadIdx.x; t < hi; t += blockDim.x) {
        if (offsets_out) offsets_out[t] = int32_t(starts[size_t(t) * stride]);
        atomicAdd(&slot[tile_class(starts, stride, t, n_tiles, n_live)], 1u);
    }
    __syncthreads();
    if (threadIdx.x < kWave) {  // exclusive scan of the class counts, 4 per lane
        constexpr uint32_t kPer = kTileClasses / kWave;
        const uint32_t lane = threadIdx.x;
        uint32_t c[kPer], sum = 0;
#pragma unroll
        for (uint32_t j = 0; j < kPer; ++j) {
            c[j] = slot[lane * kPer + j];
            sum += c[j];
        }
        uint32_t incl = sum;
#pragma unroll
        for (uint32_t off = 1; off < kWave; off <<= 1) {
            const uint32_t y = __shfl_up(incl, off, kWave);
            if (lane >= off) incl += y;
        }
        uint32_t run = incl - sum;
#pragma unroll
        for (uint32_t j = 0; j < kPer; ++j) {
            slot[lane * kPer + j] = run;
            run += c[j];
        }
    }
    __syncthreads();
    for (uint32_t t = lo + threadIdx.x; t < hi; t += blockDim.x)
        order[lo + atomicAdd(&slot[tile_class(starts, stride, t, n_tiles, n_live)], 1u)] = t;
}

// Workgroup b of a grid of 8k groups, renumbered so that the k consecutive
// ranges of the sorted list land on the 8 XCDs one each (workgroups are
// dealt to XCDs round-robin): each XCD's L2 then serves one band of cameras.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) { return (b % 8u) * (nb / 8u) + b / 8u; }

// Camera-major list k = cam * nt + t: its length in the tile-major sorted
// arrays (starts tm[t * C + cam]).
__global__ void __launch_bounds__(256)
fused_list_lengths_kernel(const uint32_t* __restrict__ tm, const uint32_t* __restrict__ d_n, uint32_t capacity,
                          uint32_t C, uint32_t nt, uint32_t* __restrict__ lengths) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= C * nt) return;
    const uint32_t n_live = live(d_n, capacity);
    const uint32_t cam = k / nt, t = k - cam * nt, j = t * C + cam;
    const uint32_t start = min(tm[j], n_live);
    const uint32_t end = (j + 1 == C * nt) ? n_live : min(tm[j + 1], n_live);
    lengths[k] = end > start ? end - start : 0u;
}

// One workgroup per camera-major list k = cam * nt + t (the lists of one XCD
// are a contiguous band of cameras): the list entries {record index,
// emission index}, the emission index w0 + (ty - y0) (x1 - x0) + (tx - x0)
// (row-major within the record's tile rectangle, the order fused_emit_kernel
// wrote it in; w0 = the record's first emission index). The lists are
// camera-major, the reference's list layout. Within each chunk of kPackChunk
// intersections the entries are ordered by the number of tile rows their
// records reach (descending; ties keep the primitive order), a stable
// counting sort in LDS: the backward's waves then take 64 records of similar
// height and loop over no more rows than their tallest record needs.
constexpr int kPackItems = 4;
constexpr int kPackThreads = 256;
constexpr int kPackChunk = kPackItems * kPackThreads;
constexpr int kPackWaves = kPackThreads / kWave;
constexpr int kPackClasses = 34;  // rows 0..32 plus "not written"
constexpr int kPackSlots = kPackItems * kPackWaves;

__global__ void __launch_bounds__(kPackThreads)
fused_pack_kernel(const uint2* __restrict__ rects, const float2* __restrict__ reach,
                  const uint32_t* __restrict__ cum, uint32_t n_recs, uint32_t C, uint32_t nt, uint32_t ntx,
                  uint32_t T, const uint32_t* __restrict__ vals, const uint32_t* __restrict__ tm,
                  const uint32_t* __restrict__ cm, const uint32_t* __restrict__ d_n, uint32_t capacity,
                  uint2* __restrict__ entries, int32_t* __restrict__ tile_offsets) {
    // cnt[class][item][wave] -> exclusive offsets in that order (class-major:
    // taller records first; within a class the chunk's own order)
    __shared__ uint32_t cnt[kPackClasses * kPackSlots];
    const uint32_t k = xcd_block(blockIdx.x, gridDim.x);
    if (k >= C * nt) return;  // uniform over the workgroup
    const uint32_t n_live = live(d_n, capacity);
    const uint32_t cam = k / nt, t = k - cam * nt;
    const uint32_t dst = cm[k], len = cm[k + 1] - dst, src = tm[t * C + cam];
    if (threadIdx.x == 0) tile_offsets[k] = int32_t(dst);
    const uint32_t ty = t / ntx, tx = t - ty * ntx;
    const float y_base = float(ty * T);
    const uint32_t wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const uint64_t below = (uint64_t(1) << lane) - 1u;
    const uint32_t skip_cls = T + 1u;
    for (uint32_t r0 = 0; r0 < len; r0 += kPackChunk) {
        uint32_t v[kPackItems], cls[kPackItems], rank[kPackItems];
#pragma unroll
        for (int q = 0; q < kPackItems; ++q) {
            const uint32_t r = r0 + q * kPackThreads + threadIdx.x;
            v[q] = (r < len && src + r < n_live) ? vals[src + r] : n_recs;
        }
        uint2 rc[kPackItems];
        uint32_t w0[kPackItems];
#pragma unroll
        for (int q = 0; q < kPackItems; ++q) {
            cls[q] = skip_cls;
            if (v[q] < n_recs) {
                rc[q] = rects[v[q]];
                w0[q] = cum[v[q]];
                const float2 e = reach[v[q]];
                cls[q] = T - rows_reached(e.x, e.y, y_base, T);
            }
        }
        for (uint32_t i = threadIdx.x; i < uint32_t(kPackClasses * kPackSlots); i += kPackThreads) cnt[i] = 0u;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kPackItems; ++q) {
            // stable rank among this wave's items of the same class: one ballot per distinct class
            uint64_t pending = __ballot(1);
            while (pending) {
                const uint32_t c = __shfl(cls[q], __builtin_ctzll(pending), kWave);
                const uint64_t m = __ballot(cls[q] == c);
                if (cls[q] == c) rank[q] = uint32_t(__popcll(m & below));
                if (lane == 0) cnt[c * kPackSlots + q * kPackWaves + wave] = uint32_t(__popcll(m));
                pending &= ~m;
            }
        }
        __syncthreads();
        if (wave == 0) {  // exclusive scan of the counts, 9 per lane
            constexpr int kPer = (kPackClasses * kPackSlots + kWave - 1) / kWave;
            uint32_t loc[kPer], sum = 0;
#pragma unroll
            for (int j = 0; j < kPer; ++j) {
                const uint32_t i = lane * kPer + j;
                loc[j] = i < uint32_t(kPackClasses * kPackSlots) ? cnt[i] : 0u;
                sum += loc[j];
            }
            uint32_t incl = sum;
#pragma unroll
            for (int off = 1; off < kWave; off <<= 1) {
                const uint32_t y = __shfl_up(incl, off, kWave);
                if (lane >= uint32_t(off)) incl += y;
            }
            uint32_t run = incl - sum;
#pragma unroll
            for (int j = 0; j < kPer; ++j) {
                const uint32_t i = lane * kPer + j;
                if (i < uint32_t(kPackClasses * kPackSlots)) cnt[i] = run;
                run += loc[j];
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kPackItems; ++q) {
            if (cls[q] == skip_cls) continue;
            const uint32_t pos = r0 + cnt[cls[q] * kPackSlots + q * kPackWaves + wave] + rank[q];
            if (dst + pos >= capacity) continue;
            const uint32_t x0 = rc[q].x & 0xffffu, x1 = rc[q].x >> 16, y0 = rc[q].y & 0xffffu;
            entries[dst + pos] = make_uint2(v[q], w0[q] + (ty - y0) * (x1 - x0) + (tx - x0));
        }
        __syncthreads();  // cnt is reused by the next chunk
    }
}

// ---------------------------------------------------------------------------
// Direct placement (the binning used when a camera's bins fit in a small LDS
// table). Bin = (tile, height class) within a camera, class 0 the tallest
// records; the packed lists are ordered (camera, tile, class, range, ...):
// each camera-major tile list holds its records by decreasing height class
// (the backward's waves take 64 records of similar height). One wave owns a
// range of R consecutive primitives of one camera, one lane per record, and
// walks the records' tile rectangles in step (pair k of every lane's
// rectangle, row-major, at step k):
//   count: per range, the number of intersections in each bin (LDS counters);
//   scan:  per (camera, bin) an exclusive scan over the ranges, then one over
//          (camera, bin): every (range, bin) gets its first packed slot;
//   place: the same walk, each (record, tile) pair's list entry {record
//          index, emission index} written straight to its slot; pairs of one
//          step that share a bin are ranked by lane (ballots on the bin's
//          bits), steps in order.
// Everything is deterministic: counts do not depend on the order of the LDS
// increments and the placement order is fixed.
// ---------------------------------------------------------------------------
constexpr uint32_t kBinClasses = 9;  // heights ceil(8 h / T) = 8 .. 1, then h = 0

__device__ __forceinline__ uint32_t height_class(const SpanRect& r, uint32_t ty, uint32_t T) {
    const uint32_t h = rows_reached(r.my, r.ext, float(ty * T), T);
    return h == 0u ? kBinClasses - 1u : kBinClasses - 1u - (8u * h + T - 1u) / T;
}

// A lane's walk over its record's tile rectangle, row-major.
struct RectWalk {
    uint32_t x0, x1, tx, ty, cnt, row, cls;
    __device__ __forceinline__ void start(const SpanRect& r, uint32_t ntx, uint32_t T) {
        x0 = r.xr & 0xffffu;
        x1 = r.xr >> 16;
        const uint32_t y0 = r.yr & 0xffffu;
        cnt = (x1 - x0) * ((r.yr >> 16) - y0);
        tx = x0;
        ty = y0;
        row = ty * ntx;
        cls = cnt ? height_class(r, ty, T) : 0u;
    }
    __device__ __forceinline__ uint32_t bin() const { return (row + tx) * kBinClasses + cls; }
    __device__ __forceinline__ void advance(const SpanRect& r, uint32_t ntx, uint32_t T) {
        if (++tx == x1) {
            tx = x0;
            ++ty;
            row += ntx;
            cls = height_class(r, ty, T);
        }
    }
};

// grid (n_ranges, C), one wave; dynamic LDS: n_bins counters.
__global__ void __launch_bounds__(kWave)
fused_bin_count_kernel(const uint2* __restrict__ rects, const float2* __restrict__ reach, uint32_t N,
                       uint32_t range_len, uint32_t n_ranges, uint32_t ntx, uint32_t T, uint32_t n_bins,
                       uint32_t* __restrict__ range_counts, uint32_t* __restrict__ counters, uint32_t capacity) {
    extern __shared__ uint32_t hist[];
    const uint32_t range = blockIdx.x, cam = blockIdx.y, lane = threadIdx.x;
    // counters = {intersections (the scan's total), overflow}: the flag is
    // raised when the count exceeds the capacity
    if (range == 0 && cam == 0 && lane == 0) counters[1] = counters[0] > capacity ? 1u : 0u;
    for (uint32_t b = lane; b < n_bins; b += kWave) hist[b] = 0u;
    __syncthreads();
    const uint32_t n0 = range * range_len, n1 = min(N, n0 + range_len);
    for (uint32_t n = n0 + lane; n < n1; n += kWave) {
        const SpanRect sr = load_span(rects, reach, size_t(cam) * N + n);
        RectWalk w;
        w.start(sr, ntx, T);
        for (uint32_t k = 0; k < w.cnt; ++k) {
            atomicAdd(&hist[w.bin()], 1u);
            w.advance(sr, ntx, T);
        }
    }
    __syncthreads();
    uint32_t* out = range_counts + (size_t(cam) * n_ranges + range) * n_bins;
    for (uint32_t b = lane; b < n_bins; b += kWave) out[b] = hist[b];
}

// Per (camera, bin): exclusive scan of its counts over the ranges (in place);
// the column total goes to totals[cam * n_bins + bin].
__global__ void __launch_bounds__(256)
fused_bin_scan_ranges_kernel(uint32_t* __restrict__ range_counts, uint32_t C, uint32_t n_ranges, uint32_t n_bins,
                             uint32_t* __restrict__ totals) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C * n_bins) return;
    const uint32_t cam = i / n_bins, b = i - cam * n_bins;
    uint32_t* col = range_counts + size_t(cam) * n_ranges * n_bins + b;
    uint32_t run = 0;
    for (uint32_t r = 0; r < n_ranges; ++r) {
        const uint32_t v = col[size_t(r) * n_bins];
        col[size_t(r) * n_bins] = run;
        run += v;
    }
    totals[i] = run;
}

// grid (n_ranges, C), one wave; dynamic LDS: the n_bins next free slots.
__global__ void __launch_bounds__(kWave)
fused_bin_place_kernel(const uint2* __restrict__ rects, const float2* __restrict__ reach,
                       const uint32_t* __restrict__ cum, uint32_t N, uint32_t range_len, uint32_t n_ranges,
                       uint32_t ntx, uint32_t T, uint32_t n_bins, uint32_t bin_bits,
                       const uint32_t* __restrict__ range_starts, const uint32_t* __restrict__ bin_starts,
                       uint32_t capacity, uint2* __restrict__ entries) {
    extern __shared__ uint32_t next_slot[];
    const uint32_t range = blockIdx.x, cam = blockIdx.y, lane = threadIdx.x;
    const uint32_t* rs = range_starts + (size_t(cam) * n_ranges + range) * n_bins;
    const uint32_t* bs = bin_starts + size_t(cam) * n_bins;
    for (uint32_t b = lane; b < n_bins; b += kWave) next_slot[b] = bs[b] + rs[b];
    __syncthreads();
    const uint64_t below = (uint64_t(1) << lane) - 1u;
    const uint32_t n0 = range * range_len, n1 = min(N, n0 + range_len);
    for (uint32_t s = n0; s < n1; s += kWave) {
        const uint32_t n = s + lane;
        const size_t idx = size_t(cam) * N + n;
        SpanRect sr = {0u, 0u, 0.f, -1.f};
        if (n < n1) sr = load_span(rects, reach, idx);
        RectWalk w;
        w.start(sr, ntx, T);
        const uint32_t g = w.cnt ? cum[idx] : 0u;
        for (uint32_t k = 0; __ballot(k < w.cnt); ++k) {
            const bool act = k < w.cnt;
            const uint32_t bin = act ? w.bin() : 0u;
            // lanes of this step with the same bin, ranked by lane
            uint64_t same = __ballot(act);
            for (uint32_t bit = 0; bit < bin_bits; ++bit) {
                const bool one = (bin >> bit) & 1u;
                const uint64_t m = __ballot(one);
                same &= one ? m : ~m;
            }
            if (act) {
                // every lane reads its bin's slot before the group's last lane advances it
                const uint32_t pos = next_slot[bin] + uint32_t(__popcll(same & below));
                if ((same >> lane) == 1u) next_slot[bin] = pos + 1u;
                if (pos < capacity) entries[pos] = make_uint2(uint32_t(idx), g + k);
                w.advance(sr, ntx, T);
            }
        }
    }
}

// Camera-major tile list starts: list (cam, t) begins at bin (cam, t, class 0).
__global__ void __launch_bounds__(256)
fused_bin_tile_offsets_kernel(const uint32_t* __restrict__ bin_starts, uint32_t n_tiles,
                              int32_t* __restrict__ tile_offsets) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n_tiles) tile_offsets[t] = int32_t(bin_starts[size_t(t) * kBinClasses]);
}

// ---------------------------------------------------------------------------
// Forward: one workgroup of WAVES waves per T x T tile (T = 8, 16, 32). Wave
// k takes batches k, k + WAVES, ... of 64 records of the tile's list: each
// lane loads one record of its next batch (coalesced vector loads issued a
// whole batch ahead, so HBM latency is hidden behind the current batch) and
// the wave reads the current batch back from its own LDS slot with broadcast
// loads. Lane = column (lane % T) and row group (lane / T), with PX = T^2 / 64
// rows per lane spaced 64 / T apart; sigma reuses the column terms across the
// lane's pixels. The waves' partial images are summed in LDS in wave order,
// so the result does not depend on scheduling.
// ---------------------------------------------------------------------------
// Adds one record to a lane's PX pixels. Bit k of the wave-uniform `rows` is
// set when row group k can hold contributing pixels; the others are skipped
// by scalar branches (one bit test each; none when every group is set).
template <int PX, bool BETA, bool VAR, typename R>
__device__ __forceinline__ void splat_accumulate(const R& s, uint32_t rows, float px, const float (&py)[PX],
                                                 float (&acc)[PX], float (&acc_var)[PX]) {
    constexpr uint32_t kAll = PX >= 32 ? 0xffffffffu : (1u << PX) - 1u;
    if (rows == 0u) return;
    const bool all = rows == kAll;
    const float dx = offset_x<BETA>(s, px);
    const float adx2 = s.A * dx * dx, bdx = s.B * dx;
#pragma unroll
    for (int k = 0; k < PX; ++k) {
        if (!all && !((rows >> k) & 1u)) continue;
        const float dy = offset_y<BETA>(s, py[k]);
        const float q = form_by_columns(adx2, bdx, s.C, dy);  // Gaussian: sigma log2 e; Beta: d^2
        float e;
        if constexpr (BETA)
            e = __builtin_amdgcn_exp2f(s.b2d * __log2f(1.f - q));
        else
            e = __builtin_amdgcn_exp2f(-q);
        const float ek = contributes(q, s.qlim) ? e : 0.f;
        if constexpr (VAR) {
            // s.unc is the staged variance weight q w^2: variance = sum q alpha^2
            acc[k] = fmaf(ek, s.w, acc[k]);
            acc_var[k] = fmaf(s.unc * ek, ek, acc_var[k]);
        } else {
            acc[k] = fmaf(ek, s.w, acc[k]);
        }
    }
}

// List entry i ({record index, emission index}), or an empty entry past the list.
__device__ __forceinline__ uint2 load_entry(const uint2* __restrict__ entries, int32_t i, int32_t hi) {
    uint2 e = make_uint2(0xffffffffu, 0u);
    if (i < hi) e = entries[i];
    return e;
}

// The record an entry names, with g = the entry's emission index (g = -1: none).
template <typename R, typename S>
__device__ __forceinline__ R gather_splat(const S* __restrict__ recs, uint2 e) {
    R s = {};
    s.g = -1;
    if (e.x != 0xffffffffu) {
        s = head(recs[e.x]);
        s.g = int32_t(e.y);
    }
    return s;
}

// Marks a record's registers as used here: the loads that produced them (a
// whole chunk earlier) are waited for at this point, before the next
// prefetch is issued, so later waits need not drain that prefetch.
template <typename R>
__device__ __forceinline__ void settle(const R& s) {
    asm volatile("" ::"v"(s.mx), "v"(s.my), "v"(s.A), "v"(s.B), "v"(s.C), "v"(s.w), "v"(s.qlim), "v"(s.g));
    if constexpr (std::is_same<R, BetaSplat>::value) asm volatile("" ::"v"(s.b2d), "v"(s.unc), "v"(s.lx), "v"(s.ly));
}

// One tile of the forward: the workgroup renders the tile at position `pos`
// of the tile order (or tile `pos` without one).
template <int T, int WAVES, bool BETA, bool VAR>
__device__ __forceinline__ void raster_fwd_tile(const uint2* __restrict__ entries,
                                                const Stored<BETA, VAR>* __restrict__ recs,
                                                const int32_t* __restrict__ offsets,
                                                const uint32_t* __restrict__ order, const uint32_t* __restrict__ d_n,
                                                uint32_t capacity, uint32_t n_tiles_total, uint32_t ntx,
                                                uint32_t nty, uint32_t W, uint32_t H, float* __restrict__ out,
                                                float* __restrict__ out_var, uint32_t pos) {
    constexpr int PX = T * T / kWave, STEP = kWave / T;
    static_assert(PX * kWave == T * T && STEP * T == kWave, "tile must be 8, 16 or 32");
    __shared__ float part[WAVES][T * T];  // the slices' partial images (mean, then variance)
    using R = Rec<BETA, VAR>;
    __shared__ R stage[WAVES][kWave];
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const uint32_t tile = order ? order[pos] : pos;
    const uint32_t n_live = live(d_n, capacity);
    int32_t start, end;
    fused_tile(offsets, tile, n_tiles_total, n_live, start, end);
    start = __builtin_amdgcn_readfirstlane(start);  // uniform: keep the loop bounds scalar
    end = __builtin_amdgcn_readfirstlane(end);
    // wave k takes batches k, k + WAVES, ... of 64 records (the lists are ordered
    // by height class, so contiguous slices would give the first waves the tallest)
    constexpr int32_t kStride = WAVES * kWave;
    const int32_t lo = start + int32_t(wave) * kWave, hi = end;
    const uint32_t nt = ntx * nty, cam = tile / nt, t = tile - cam * nt;
    const uint32_t ty = t / ntx, tx = t - ty * ntx;
    const uint32_t lane = threadIdx.x % kWave, col = lane % T, row0 = lane / T;
    const float px = float(tx * T + col) + 0.5f;
    float py[PX], acc[PX], acc_var[PX];
#pragma unroll
    for (int k = 0; k < PX; ++k) {
        py[k] = float(ty * T + row0 + STEP * k) + 0.5f;
        acc[k] = 0.f;
        acc_var[k] = 0.f;
    }
    R* slot = stage[wave];
    // two-stage prefetch: a batch's records are gathered one batch ahead, their
    // list entries two batches ahead
    uint2 ahead = load_entry(entries, lo + int32_t(lane), hi);
    R next = gather_splat<R>(recs, ahead);
    ahead = load_entry(entries, lo + kStride + int32_t(lane), hi);
    const float y_base = float(ty * T);
    for (int32_t base = lo; base < hi; base += kStride) {
        // row groups of the record this lane stages (lane rows are row0 + STEP k,
        // row0 < STEP); record j's span is read from lane j's register
        int r_lo, r_hi;
        row_span(next, y_base, T, r_lo, r_hi);
        const int k_lo = r_lo / STEP, k_hi = (r_hi + STEP - 1) / STEP;  // groups [k_lo, k_hi)
        const uint32_t spans = r_lo < r_hi ? ((k_hi >= 32 ? 0xffffffffu : (1u << k_hi) - 1u) & ~((1u << k_lo) - 1u))
                                           : 0u;
        wave_lds_sync();  // every lane is done reading the previous batch
        if constexpr (VAR) {
            // the staged copy's uncertainty field carries the record's variance
            // weight q w^2 (variance = sum q alpha^2 = sum (q w^2) e^2)
            R staged = next;
            staged.unc = next.unc * next.w * next.w;
            slot[lane] = staged;
        } else {
            slot[lane] = next;
        }
        wave_lds_sync();
        next = gather_splat<R>(recs, ahead);
        ahead = load_entry(entries, base + 2 * kStride + int32_t(lane), hi);
        const int32_t n = min(int32_t(kWave), hi - base);
        // LDS reads pipelined one record ahead in two registers that alternate
        // roles (no copies between them); the batch holds n >= 1 records
        // (settle waits for the current record before the next read is issued, so
        // that read overlaps the current record's work)
        R a = slot[0];
        for (int32_t j = 0;;) {
            settle(a);
            const R b = slot[min(j + 1, int32_t(kWave) - 1)];
            splat_accumulate<PX, BETA, VAR>(a, __builtin_amdgcn_readlane(spans, j), px, py, acc, acc_var);
            if (++j >= n) break;
            settle(b);
            a = slot[min(j + 1, int32_t(kWave) - 1)];
            splat_accumulate<PX, BETA, VAR>(b, __builtin_amdgcn_readlane(spans, j), px, py, acc, acc_var);
            if (++j >= n) break;
        }
    }
#pragma unroll
    for (int k = 0; k < PX; ++k) part[wave][(row0 + STEP * k) * T + col] = acc[k];
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < uint32_t(T * T); p += WAVES * kWave) {
        float v = 0.f;
#pragma unroll
        for (int k = 0; k < WAVES; ++k) v += part[k][p];
        const uint32_t x = tx * T + p % T, y = ty * T + p / T;
        if (x < W && y < H) out[(size_t(cam) * H + y) * W + x] = v;
    }
    if constexpr (VAR) {
        __syncthreads();  // the mean's partial images are consumed
#pragma unroll
        for (int k = 0; k < PX; ++k) part[wave][(row0 + STEP * k) * T + col] = acc_var[k];
        __syncthreads();
        for (uint32_t p = threadIdx.x; p < uint32_t(T * T); p += WAVES * kWave) {
            float v = 0.f;
#pragma unroll
            for (int k = 0; k < WAVES; ++k) v += part[k][p];
            const uint32_t x = tx * T + p % T, y = ty * T + p / T;
            if (x < W && y < H) out_var[(size_t(cam) * H + y) * W + x] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// Backward: one workgroup of WAVES waves per T x T tile; wave k takes chunks
// k, k + WAVES, ... of 64 records, one lane per record. The tile's adjoint
// image sits in LDS and is read row by row; each lane walks only the rows its
// record can reach (row_span; the lists order records by height class, so a
// chunk's records reach similar numbers of rows) and sums its own primitive's
// moments over them
// without any cross-lane reduction:
//   g = v e,  M0 = sum g, Mx = sum g dx, My = sum g dy, Mxx, Mxy, Myy
// with dx = px - mx (the moments are summed directly in dx: expanding them
// around the tile corner cancels badly for primitives centred in the tile).
// The VJP is
//   v_mean = w (a Mx + b My, b Mx + c My), v_conic = -w (Mxx/2, Mxy, Myy/2),
//   v_w = M0, stored (no atomics) as the partial of emission index g.
// Beta: with h = 1 - d^2 the moments are of G = v h^(b2d - 1), plus
// Se = sum v h^b2d and Sl = sum v h^b2d log2 h, and (delta = mean - pixel)
//   v_mean = w b2d (2a Mx + 2b My, 2b Mx + 2c My),
//   v_conic = -w b2d (Mxx, 2 Mxy, Myy), v_w = Se, v_b2d = w ln2 Sl.
// With an uncertainty q (variance = sum q alpha^2) the pixel adjoint becomes
// v + 2 q alpha v_var (v alone when the variance is detached) and the partial
// also carries v_q = sum v_var alpha^2 (its eighth slot).
// ---------------------------------------------------------------------------
// Waves per SIMD the backward asks the register allocator for: the Gaussian
// 16-pixel-tile kernel fits 72 VGPRs without spills, 7 waves per SIMD instead
// of 6 (its 75-VGPR default); the other variants keep the default (forcing
// the next occupancy step spills them).
template <int T, bool BETA, bool VAR>
constexpr int bwd_waves_per_eu() { return (T == 16 && !BETA && !VAR) ? 5 : 1; }

// One tile of the backward (the tile at position `pos` of the tile order).
template <int T, int WAVES, bool BETA, bool VAR>
__device__ __forceinline__ void raster_bwd_tile(const uint2* __restrict__ entries,
                                                const Stored<BETA, VAR>* __restrict__ recs,
                                                const int32_t* __restrict__ offsets,
                                                const uint32_t* __restrict__ order, const uint32_t* __restrict__ d_n,
                                                uint32_t capacity, uint32_t n_tiles_total, uint32_t ntx,
                                                uint32_t nty, uint32_t W, uint32_t H, const float* __restrict__ v_img,
                                                const float* __restrict__ v_var, int detach,
                                                float* __restrict__ isect_grads, uint32_t pos) {
    __shared__ __align__(16) float sv[VAR ? 2 : 1][T * T];
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const uint32_t tile = order ? order[pos] : pos;
    const uint32_t n_live = live(d_n, capacity);
    int32_t start, end;
    fused_tile(offsets, tile, n_tiles_total, n_live, start, end);
    if (start == end) return;  // uniform over the workgroup
    const uint32_t nt = ntx * nty, cam = tile / nt, t = tile - cam * nt;
    const uint32_t ty = t / ntx, tx = t - ty * ntx;
    const uint32_t lane = threadIdx.x % kWave;
    for (uint32_t p = threadIdx.x; p < uint32_t(T * T); p += WAVES * kWave) {
        const uint32_t x = tx * T + p % T, y = ty * T + p / T;
        const bool in = x < W && y < H;
        const size_t o = (size_t(cam) * H + y) * W + x;
        sv[0][p] = in ? v_img[o] : 0.f;
        if constexpr (VAR) sv[VAR ? 1 : 0][p] = (in && v_var) ? v_var[o] : 0.f;
    }
    __syncthreads();
    // the variance adjoint reaches alpha (2 q alpha v_var) unless the variance is detached
    const float var_scale = (VAR && !detach) ? 2.f : 0.f;
    const float x0 = float(tx * T) + 0.5f, y0 = float(ty * T) + 0.5f;
    constexpr int32_t kStride = WAVES * kWave;
    using R = Rec<BETA, VAR>;
    int32_t chunk = start + int32_t(wave) * kWave;
    // records gathered a chunk ahead, list entries two chunks ahead
    uint2 ahead = load_entry(entries, chunk + int32_t(lane), end);
    R next = gather_splat<R>(recs, ahead);
    uint32_t next_rec = ahead.x;
    ahead = load_entry(entries, chunk + kStride + int32_t(lane), end);
    for (; chunk < end; chunk += kStride) {
        const int32_t i = chunk + int32_t(lane);
        const bool active = i < end;
        const R s = next;
        const uint32_t rec = next_rec;  // Beta: the record whose double-precision form settles edge pairs
        settle(s);
        next = gather_splat<R>(recs, ahead);
        next_rec = ahead.x;
        ahead = load_entry(entries, chunk + 2 * kStride + int32_t(lane), end);
        const float dx0 = offset_x<BETA>(s, x0);
        float m0 = 0.f, mx = 0.f, mxx = 0.f, my = 0.f, mxy = 0.f, myy = 0.f, se = 0.f, sl = 0.f, vq = 0.f;
        float bm1 = 0.f, kvar = 0.f;
        if constexpr (BETA) bm1 = s.b2d - 1.f;
        if constexpr (VAR) kvar = var_scale * s.unc * s.w;  // d variance / d (alpha / w) factor
        // only the rows this lane's record can reach (the others add exact zeros);
        // the wave runs as many rows as its tallest record needs
        int r_lo = 0, r_hi = 0;
        if (active) row_span(s, y0 - 0.5f, T, r_lo, r_hi);
        // Beta: pairs below lim.x pass; those in [lim.x, lim.y) are settled in double
        const uint2 lim = (BETA && active) ? edge_limits(s) : make_uint2(0u, 0u);
        // the columns' offsets and their squares are the same in every row
        float dxc[T], dxsq[T];
#pragma unroll
        for (int j = 0; j < T; ++j) {
            dxc[j] = dx0 + float(j);
            dxsq[j] = dxc[j] * dxc[j];
        }
        for (int r = r_lo; r < r_hi; ++r) {
            const float dy = offset_y<BETA>(s, y0 + float(r));
            const float bdy = s.B * dy, cdy2 = s.C * dy * dy;
            const float4* row = reinterpret_cast<const float4*>(sv[0] + r * T);
            float v[T], vv[VAR ? T : 1];
#pragma unroll
            for (int k = 0; k < T / 4; ++k) {
                const float4 q4 = row[k];
                v[4 * k] = q4.x;
                v[4 * k + 1] = q4.y;
                v[4 * k + 2] = q4.z;
                v[4 * k + 3] = q4.w;
            }
            if constexpr (VAR) {
                const float4* rowv = reinterpret_cast<const float4*>(sv[VAR ? 1 : 0] + r * T);
#pragma unroll
                for (int k = 0; k < T / 4; ++k) {
                    const float4 q4 = rowv[k];
                    vv[4 * k] = q4.x;
                    vv[4 * k + 1] = q4.y;
                    vv[4 * k + 2] = q4.z;
                    vv[4 * k + 3] = q4.w;
                }
            }
            float r0 = 0.f, r1 = 0.f, r2 = 0.f;
            // one Beta pair's terms (in: it passes the cut-off). h is clamped to
            // the smallest normal float (outside the support 1 - q <= 0), so log2 h
            // is finite and a pair that fails the cut-off (hb = 0, so ve = 0) adds
            // exact zeros to every sum
            auto beta_pair = [&](float q, float dx, float dxsq_j, float vi, float vvi, bool in) {
                const float h = fmaxf(1.f - q, 1.17549435e-38f), lg = __log2f(h);
                const float hb = in ? __builtin_amdgcn_exp2f(bm1 * lg) : 0.f;  // h^(b2d - 1)
                float g, ve;
                if constexpr (VAR) {
                    const float he = hb * h;  // h^b2d = alpha / w
                    const float t = vvi * he;
                    const float vj = fmaf(kvar, t, vi);
                    g = vj * hb;
                    ve = vj * he;
                    vq = fmaf(t, he, vq);
                } else {
                    g = vi * hb;
                    ve = g * h;
                }
                se += ve;
                sl = fmaf(ve, lg, sl);
                r0 += g;
                r1 = fmaf(g, dx, r1);
                r2 = fmaf(g, dxsq_j, r2);
            };
            bool any_near = false;  // Beta: a pair of this row lies in the edge band
#pragma unroll
            for (int j = 0; j < T; ++j) {
                const float dx = dxc[j];
                const float q = form_by_rows(s.A, dx, bdy, cdy2);
                if constexpr (BETA) {
                    const bool in = contributes(q, lim.x);
                    any_near = any_near || (!in && contributes(q, lim.y));
                    beta_pair(q, dx, dxsq[j], v[j], VAR ? vv[VAR ? j : 0] : 0.f, in);
                } else {
                    const float e = contributes(q, s.qlim) ? __builtin_amdgcn_exp2f(-q) : 0.f;
                    float g;
                    if constexpr (VAR) {
                        const float t = vv[j] * e;
                        const float vj = fmaf(kvar, t, v[j]);
                        g = vj * e;
                        vq = fmaf(t, e, vq);
                    } else {
                        g = e * v[j];
                    }
                    r0 += g;
                    r1 = fmaf(g, dx, r1);
                    r2 = fmaf(g, dxsq[j], r2);
                }
            }
            if constexpr (BETA) {
                // the pairs of this row in the edge band, settled in double (rare)
                if (__builtin_expect(__ballot(any_near) != 0ull, 0) && any_near) {
#pragma unroll 1
                    for (int j = 0; j < T; ++j) {
                        const float dx = dx0 + float(j);
                        const float q = form_by_rows(s.A, dx, bdy, cdy2);
                        if (!contributes(q, lim.x) && contributes(q, lim.y) &&
                            edge_keep(recs[rec].e, x0 + float(j), y0 + float(r)))
                            beta_pair(q, dx, dx * dx, sv[0][r * T + j], VAR ? sv[VAR ? 1 : 0][r * T + j] : 0.f, true);
                    }
                }
            }
            m0 += r0;
            mx += r1;
            mxx += r2;
            my = fmaf(dy, r0, my);
            mxy = fmaf(dy, r1, mxy);
            myy = fmaf(dy * dy, r0, myy);
        }
        if (active && s.g >= 0 && uint32_t(s.g) < n_live) {
            float* out = isect_grads + kPartialStride * size_t(s.g);
            const float wq = s.w * s.w * vq;  // sum v_var alpha^2 (the uncertainty's gradient)
            if constexpr (BETA) {
                const float a = s.A, b = 0.5f * s.B, c = s.C, k = s.w * s.b2d;
                float4* gp = reinterpret_cast<float4*>(out);
                gp[0] = make_float4(k * (2.f * a * mx + 2.f * b * my), k * (2.f * b * mx + 2.f * c * my), -k * mxx,
                                    -2.f * k * mxy);
                gp[1] = make_float4(-k * myy, se, s.w * 0.6931471805599453f * sl, wq);
            } else {
                // conic entries in pixel units: a = 2 A / log2 e, b = B / log2 e, c = 2 C / log2 e
                const float a = 2.f * s.A / kLog2e, b = s.B / kLog2e, c = 2.f * s.C / kLog2e;
                const float g0 = s.w * (a * mx + b * my), g1 = s.w * (b * mx + c * my), g2 = -0.5f * s.w * mxx,
                            g3 = -s.w * mxy, g4 = -0.5f * s.w * myy;
                float4* gp = reinterpret_cast<float4*>(out);
                gp[0] = make_float4(g0, g1, g2, g3);
                gp[1] = make_float4(g4, m0, 0.f, VAR ? wq : 0.f);
            }
        }
    }
}

// One job: tile at the XCD-banded position of this workgroup.
template <int T, int WAVES, bool BETA, bool VAR>
__global__ void __launch_bounds__(WAVES * kWave)
fused_raster_fwd_kernel(const uint2* __restrict__ entries, const Stored<BETA, VAR>* __restrict__ recs,
                        const int32_t* __restrict__ offsets,
                        const uint32_t* __restrict__ order, const uint32_t* __restrict__ d_n,
                        uint32_t capacity, uint32_t n_tiles_total, uint32_t ntx, uint32_t nty, uint32_t W,
                        uint32_t H, float* __restrict__ out, float* __restrict__ out_var) {
    raster_fwd_tile<T, WAVES, BETA, VAR>(entries, recs, offsets, order, d_n, capacity, n_tiles_total, ntx, nty, W, H,
                                         out, out_var, order ? xcd_slot(blockIdx.x, gridDim.x) : blockIdx.x);
}

template <int T, int WAVES, bool BETA, bool VAR>
__global__ void __launch_bounds__(WAVES * kWave) __attribute__((amdgpu_waves_per_eu(bwd_waves_per_eu<T, BETA, VAR>())))
fused_raster_bwd_kernel(const uint2* __restrict__ entries, const Stored<BETA, VAR>* __restrict__ recs,
                        const int32_t* __restrict__ offsets,
                        const uint32_t* __restrict__ order, const uint32_t* __restrict__ d_n,
                        uint32_t capacity, uint32_t n_tiles_total, uint32_t ntx, uint32_t nty, uint32_t W,
                        uint32_t H, const float* __restrict__ v_img, const float* __restrict__ v_var, int detach,
                        float* __restrict__ isect_grads) {
    raster_bwd_tile<T, WAVES, BETA, VAR>(entries, recs, offsets, order, d_n, capacity, n_tiles_total, ntx, nty, W, H,
                                         v_img, v_var, detach, isect_grads,
                                         order ? xcd_slot(blockIdx.x, gridDim.x) : blockIdx.x);
}

// ---------------------------------------------------------------------------
// Packed launches: the tiles of several jobs (independent reconstructions of
// any size with the same tile edge, waves and record kind) in one launch.
// Job j's n_j tiles are jobs[j].first_tile .. first_tile + n_j - 1 of the
// concatenation. Each XCD x takes band x of every job (the band its own launch
// would give XCD x: per_j + (x < rem_j) positions from x per_j + min(x,
// rem_j), per_j = n_j / 8, rem_j = n_j % 8), jobs in table order, so the
// per-tile cost of different jobs (which differs by orders of magnitude)
// spreads evenly over the XCDs. Workgroup b (XCD b % 8) takes item b / 8 of
// its XCD's sequence; the grid is 8 x the longest sequence and workgroups past
// the end of theirs exit. The job is found by a scan of the table
// (wave-uniform: scalar loads).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t packed_band(uint32_t n, uint32_t x) { return n / 8u + (x < n % 8u ? 1u : 0u); }

// The job of workgroup b (n_jobs: none) and its position in that job's tile order.
__device__ __forceinline__ uint32_t packed_job(const r2s_raster_job* __restrict__ jobs, uint32_t n_jobs,
                                               uint32_t total, uint32_t b, uint32_t& pos) {
    const uint32_t x = b % 8u;
    uint32_t k = b / 8u;
    for (uint32_t j = 0; j < n_jobs; ++j) {
        const uint32_t n = (j + 1 < n_jobs ? jobs[j + 1].first_tile : total) - jobs[j].first_tile;
        const uint32_t size = packed_band(n, x);
        if (k < size) {
            pos = x * (n / 8u) + min(x, n % 8u) + k;
            return j;
        }
        k -= size;
    }
    return n_jobs;
}

template <int T>
__device__ __forceinline__ void packed_grid(const r2s_raster_job& J, uint32_t& ntx, uint32_t& nty, uint32_t& nt) {
    ntx = (J.width + T - 1) / T;
    nty = (J.height + T - 1) / T;
    nt = J.C * ntx * nty;
}

template <int T, int WAVES, bool BETA, bool VAR>
__global__ void __launch_bounds__(WAVES * kWave)
fused_raster_fwd_packed_kernel(const r2s_raster_job* __restrict__ jobs, uint32_t n_jobs, uint32_t total) {
    uint32_t pos = 0;
    const uint32_t j = packed_job(jobs, n_jobs, total, blockIdx.x, pos);
    if (j == n_jobs) return;  // past the end of this XCD's sequence (uniform over the workgroup)
    const r2s_raster_job& J = jobs[j];
    uint32_t ntx, nty, nt;
    packed_grid<T>(J, ntx, nty, nt);
    const auto* recs = reinterpret_cast<const Stored<BETA, VAR>*>(static_cast<const char*>(J.packed) +
                                                                   entries_bytes(J.capacity));
    raster_fwd_tile<T, WAVES, BETA, VAR>(static_cast<const uint2*>(J.packed), recs, J.tile_offsets, J.tile_order,
                                         J.counters, J.capacity, nt, ntx, nty, J.width, J.height, J.image,
                                         J.image_var, pos);
}

template <int T, int WAVES, bool BETA, bool VAR>
__global__ void __launch_bounds__(WAVES * kWave) __attribute__((amdgpu_waves_per_eu(bwd_waves_per_eu<T, BETA, VAR>())))
fused_raster_bwd_packed_kernel(const r2s_raster_job* __restrict__ jobs, uint32_t n_jobs, uint32_t total) {
    uint32_t pos = 0;
    const uint32_t j = packed_job(jobs, n_jobs, total, blockIdx.x, pos);
    if (j == n_jobs) return;
    const r2s_raster_job& J = jobs[j];
    uint32_t ntx, nty, nt;
    packed_grid<T>(J, ntx, nty, nt);
    const auto* recs = reinterpret_cast<const Stored<BETA, VAR>*>(static_cast<const char*>(J.packed) +
                                                                   entries_bytes(J.capacity));
    raster_bwd_tile<T, WAVES, BETA, VAR>(static_cast<const uint2*>(J.packed), recs, J.tile_offsets, J.tile_order,
                                         J.counters, J.capacity, nt, ntx, nty, J.width, J.height, J.v_image,
                                         J.v_image_var, J.detach, J.isect_grads, pos);
}

// ---------------------------------------------------------------------------
// Keep masks (diagnostic): which (list entry, pixel) pairs pass the cut-off in
// the forward and in the backward, decided by the same operations as those
// kernels (row_span, their row-group / row skipping, form_by_columns /
// form_by_rows, contributes). One workgroup per tile; one thread per
// (entry, 32-pixel word). Bit b of word w of an entry is pixel p = 32 w + b =
// row * T + column of the entry's tile. Tests evaluate the kept pairs in
// float64 to separate arithmetic error from pairs at the cut-off.
// ---------------------------------------------------------------------------
template <int T, bool BETA, bool VAR>
__global__ void __launch_bounds__(256)
fused_keep_mask_kernel(const uint2* __restrict__ entries, const Stored<BETA, VAR>* __restrict__ recs,
                       const int32_t* __restrict__ offsets, const uint32_t* __restrict__ d_n, uint32_t capacity,
                       uint32_t n_tiles_total, uint32_t ntx, uint32_t nty, uint32_t* __restrict__ keep_fwd,
                       uint32_t* __restrict__ keep_bwd) {
    constexpr int WORDS = T * T / 32, STEP = kWave / T;
    const uint32_t tile = blockIdx.x;
    const uint32_t n_live = live(d_n, capacity);
    int32_t start, end;
    fused_tile(offsets, tile, n_tiles_total, n_live, start, end);
    const uint32_t nt = ntx * nty, t = tile - (tile / nt) * nt;
    const uint32_t ty = t / ntx, tx = t - ty * ntx;
    const float y_base = float(ty * T);
    const float x0 = float(tx * T) + 0.5f, y0 = float(ty * T) + 0.5f;
    for (int64_t i = int64_t(start) * WORDS + threadIdx.x; i < int64_t(end) * WORDS; i += blockDim.x) {
        const int32_t entry = int32_t(i / WORDS), word = int32_t(i % WORDS);
        const uint32_t rec = entries[entry].x;
        const Rec<BETA, VAR> s = head(recs[rec]);
        int r_lo, r_hi;
        row_span(s, y_base, T, r_lo, r_hi);
        const int k_lo = r_lo / STEP, k_hi = (r_hi + STEP - 1) / STEP;  // the forward's row groups
        // the backward's decision: the float32 cut-off; Beta: below the edge band
        // pass, in it settled in double (the forward takes the float32 cut-off)
        const uint2 lim = BETA ? edge_limits(s) : make_uint2(0u, 0u);
        auto keeps_bwd = [&](float q, float px, float py) {
            if constexpr (BETA) {
                if (contributes(q, lim.x)) return true;
                return contributes(q, lim.y) && edge_keep(recs[rec].e, px, py);
            }
            return contributes(q, s.qlim);
        };
        uint32_t bits_fwd = 0u, bits_bwd = 0u;
        for (int b = 0; b < 32; ++b) {
            const int pix = word * 32 + b, r = pix / T, col = pix % T;
            const float px = float(tx * T + uint32_t(col)) + 0.5f, py = float(ty * T + uint32_t(r)) + 0.5f;
            if (r_lo < r_hi && r / STEP >= k_lo && r / STEP < k_hi) {
                const float dx = offset_x<BETA>(s, px);
                const float adx2 = s.A * dx * dx, bdx = s.B * dx;
                const float dy = offset_y<BETA>(s, py);
                if (contributes(form_by_columns(adx2, bdx, s.C, dy), s.qlim)) bits_fwd |= 1u << b;
            }
            if (r >= r_lo && r < r_hi) {
                const float dx = offset_x<BETA>(s, x0) + float(col);
                const float dy = offset_y<BETA>(s, y0 + float(r));
                const float bdy = s.B * dy, cdy2 = s.C * dy * dy;
                if (keeps_bwd(form_by_rows(s.A, dx, bdy, cdy2), px, py)) bits_bwd |= 1u << b;
            }
        }
        keep_fwd[i] = bits_fwd;
        keep_bwd[i] = bits_bwd;
    }
}

// ---------------------------------------------------------------------------
// Parameter VJP, one pass. A workgroup takes 64 primitives (one per lane) and
// kVjpWaves camera phases (wave k takes cameras k, k + kVjpWaves, ...): per
// (camera, primitive) the record's intersection partials [cum[v], cum[v+1])
// are summed in emission order, then the conic/mu VJP (double) and the
// projection VJP give that camera's contribution to v_mean (3),
// v_covariance (6), v_density (1), v_beta (1) and v_uncertainty (1, the
// partials' eighth slot). Each thread sums its
// cameras in order, the waves' sums are added in LDS in wave order, and the
// parameter-transform VJP finishes each primitive: a fixed order throughout.
// A partial is (v_mx, v_my, v_a, v_b, v_c, v_w[, v_b2d]) with w = density *
// mu [* ratio(beta)]; for Beta primitives v_beta = v_b2d + v_mus mus dlog
// ratio/dbeta and the conic/mu VJP receives v_mu = v_mus ratio
// (r2splat/explicit.py:227-231).
// ---------------------------------------------------------------------------
constexpr int kParamFields = 12;  // v_mean 3, v_covariance 6, v_density, v_beta, v_uncertainty
constexpr int kVjpWaves = 8;  // 8 waves fit the kernel in registers without spills
// partials one wave stages at a time (8 floats each): 192 (a camera's share of
// a 64-primitive group is about 275 partials at 100k Gaussians x 64 views)
// takes 0.30 ms where 96 took 0.33 ms per parameter VJP; the LDS stays within
// the two workgroups per CU the registers allow
constexpr uint32_t kVjpStageParts = 192;
// one wave's share of the LDS buffer: its staged partials, later its sums
constexpr uint32_t kVjpWaveFloats = kVjpStageParts * kPartialStride;
static_assert(kVjpWaveFloats >= uint32_t(kParamFields * kWave), "a wave's sums fit its staging area");

template <bool BETA, bool VAR>
__global__ void __launch_bounds__(kVjpWaves * kWave)
fused_param_vjp_kernel(const float* __restrict__ means, const float* __restrict__ quats,
                       const float* __restrict__ scales, const float* __restrict__ dens,
                       const float* __restrict__ ratio, const float* __restrict__ log_grad,
                       const float* __restrict__ viewmats, const float* __restrict__ Ks, uint32_t W, uint32_t H,
                       float eps, uint32_t C, uint32_t N, const uint32_t* __restrict__ cum,
                       const uint32_t* __restrict__ d_n, uint32_t capacity, const float* __restrict__ isect_grads,
                       float* __restrict__ v_means, float* __restrict__ v_quats, float* __restrict__ v_scales,
                       float* __restrict__ v_dens, float* __restrict__ v_betas, float* __restrict__ v_unc,
                       const uint32_t* __restrict__ perm) {
    // one LDS buffer: each wave's staging area during the camera loop, then the
    // same wave's partial sums (part[wave][field][lane]) for the final reduction
    __shared__ __align__(16) float lds[kVjpWaves * kVjpWaveFloats];
    const uint32_t lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    // lane p of the primitive order: records (c, p), parameters and gradients of input primitive n
    const uint32_t pr = blockIdx.x * kWave + lane;
    const bool valid = pr < N;
    const uint32_t n = valid ? perm[pr] : pr;
    const uint32_t last = min(N - blockIdx.x * kWave, uint32_t(kWave)) - 1u;  // the last valid lane
    float* stage = lds + wave * kVjpWaveFloats;
    float acc[kParamFields] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float cov[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, p[3] = {0.f, 0.f, 0.f};
    const uint32_t n_live = live(d_n, capacity);
    if (valid) {
        // the rotation and scales are recomputed for the final transform VJP
        // rather than kept live across the camera loop
        float q[4], mag, norm, R[9];
        unit_quat(quats + 4 * size_t(n), q, mag, norm);
        quat_to_rot(q, R);
        const float s[3] = {scales[3 * size_t(n)], scales[3 * size_t(n) + 1], scales[3 * size_t(n) + 2]};
        covariance_from(R, s, cov);
        p[0] = means[3 * size_t(n)];
        p[1] = means[3 * size_t(n) + 1];
        p[2] = means[3 * size_t(n) + 2];
    }
    {
        for (uint32_t c = wave; c < C; c += kVjpWaves) {
            const size_t v = size_t(c) * N + pr;
            uint32_t lo = 0, hi = 0;
            if (valid) {
                lo = min(cum[v], n_live);
                hi = min(cum[v + 1], n_live);
            }
            // the wave's 64 positions own one contiguous range of partials in this
            // camera: it is staged through LDS with coalesced loads, then each lane
            // sums its own records' partials in emission order
            const uint32_t wlo = __builtin_amdgcn_readfirstlane(lo);
            const uint32_t whi = __builtin_amdgcn_readlane(hi, last);
            float g0 = 0.f, g1 = 0.f, g2 = 0.f, g3 = 0.f, g4 = 0.f, g5 = 0.f, g6 = 0.f, g7 = 0.f;
            for (uint32_t base = wlo; base < whi; base += kVjpStageParts) {
                const uint32_t cnt = min(kVjpStageParts, whi - base);
                wave_lds_sync();  // every lane is done with the previous chunk
                for (uint32_t i = lane; i < cnt; i += kWave) {
                    const float4* gp = reinterpret_cast<const float4*>(isect_grads + kPartialStride * size_t(base + i));
                    float4* sp = reinterpret_cast<float4*>(stage + kPartialStride * i);
                    sp[0] = gp[0];
                    sp[1] = gp[1];
                }
                wave_lds_sync();
                const uint32_t j1 = min(hi, base + cnt);
                for (uint32_t j = max(lo, base); j < j1; ++j) {
                    const float4* gp = reinterpret_cast<const float4*>(stage + kPartialStride * (j - base));
                    const float4 a = gp[0], b = gp[1];
                    g0 += a.x; g1 += a.y; g2 += a.z; g3 += a.w;
                    g4 += b.x; g5 += b.y;
                    if constexpr (BETA || VAR) {
                        g6 += b.z;
                        g7 += b.w;
                    }
                }
            }
            if (!valid || (g0 == 0.f && g1 == 0.f && g2 == 0.f && g3 == 0.f && g4 == 0.f && g5 == 0.f &&
                           g6 == 0.f && g7 == 0.f))
                continue;
            acc[11] += g7;  // the uncertainty enters the variance directly
            Camera<float> cam;
            cam.load(viewmats + 16 * size_t(c), Ks + 9 * size_t(c));
            ProjState<float> st;
            if (!project_state(cam, p, cov, float(W), float(H), st)) continue;
            float c2[6];
            projected_cov(st, c2);
            double cd[6], P[3], mu;
#pragma unroll
            for (int k = 0; k < 6; ++k) cd[k] = double(c2[k]);
            conic_mu_fast_f64(cd, double(eps), P, mu);
            // the forward stored the conic and mu in float32
            const double Pf[3] = {double(float(P[0])), double(float(P[1])), double(float(P[2]))};
            const double muf = double(float(mu));
            const double vP[3] = {double(g2), double(g3), double(g4)};
            float out[kParamFields - 1] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            double v_mu;
            if constexpr (BETA) {
                const float mus = float(mu) * ratio[n];
                const float v_mus = g5 * dens[n];
                out[9] = g5 * mus;
                out[10] = g6 + v_mus * mus * log_grad[n];
                v_mu = double(v_mus) * double(ratio[n]);
            } else {
                out[9] = float(muf) * g5;
                v_mu = double(dens[n]) * double(g5);
            }
            double gcov[6];
            conic_mu_vjp_f64(cd, Pf, muf, vP, v_mu, gcov);
            const float vm[2] = {g0, g1};
            float vc[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) vc[k] = float(gcov[k]);
            projection_vjp(cam, st, vm, vc, out, out + 3);
#pragma unroll
            for (int f = 0; f < kParamFields - 1; ++f) acc[f] += out[f];
        }
    }
    wave_lds_sync();  // the wave's last chunk is read before its sums overwrite the area
#pragma unroll
    for (int f = 0; f < kParamFields; ++f) stage[f * kWave + lane] = acc[f];
    __syncthreads();
    if (wave != 0 || pr >= N) return;
#pragma unroll
    for (int f = 0; f < kParamFields; ++f) {
        float t = lds[f * kWave + lane];
        for (int w = 1; w < kVjpWaves; ++w) t += lds[w * kVjpWaveFloats + f * kWave + lane];
        acc[f] = t;
    }
    float q[4], mag, norm, R[9];
    unit_quat(quats + 4 * size_t(n), q, mag, norm);
    quat_to_rot(q, R);
    const float s[3] = {scales[3 * size_t(n)], scales[3 * size_t(n) + 1], scales[3 * size_t(n) + 2]};
    float gq[4], vs[3];
    covariance_vjp(q, R, s, acc + 3, gq, vs);
    normalize_vjp(q, mag, norm, gq, v_quats + 4 * size_t(n));
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        v_means[3 * size_t(n) + k] = acc[k];
        v_scales[3 * size_t(n) + k] = vs[k];
    }
    v_dens[n] = acc[9];
    if (v_betas) v_betas[n] = acc[10];
    if (v_unc) v_unc[n] = acc[11];
}

// ---------------------------------------------------------------------------
// Forward scratch layout (released by the caller after the forward)
// ---------------------------------------------------------------------------
struct FusedScratch {
    size_t rects, scan, sort, sort_bytes, keys, vals, tm, lengths, cm,
        tile_scan, bytes;
    // direct placement (place = true) or sort + pack
    bool place;
    uint32_t range_len, n_ranges, n_bins, bin_bits;
    size_t reach, range_counts, bin_totals, bin_scan;
    // primitive order: bounds, key sort, and the permuted parameter copies
    size_t order_bounds, order_sort_n, order_sort_n_bytes, order_sorted_keys, params, frames, cams;
};

// floats per primitive of the permuted copies: mean 3, density, beta, ratio,
// uncertainty (then a PrimFrame per primitive)
constexpr size_t kPermutedFloats = 7;

static size_t al(size_t x) { return (x + 255) & ~size_t(255); }

// kind: bit 0 = Beta primitives, bit 1 = primitives with an uncertainty
// (variance output): 32-byte records, 48 with an uncertainty, 96 for Beta.
constexpr int kKindBeta = 1, kKindVar = 2;
static size_t record_bytes(int kind) {
    return (kind & kKindBeta) ? sizeof(BetaRec) : (kind & kKindVar) ? sizeof(BetaSplat) : sizeof(Splat);
}

static size_t packed_bytes(uint32_t C, uint32_t N, uint32_t capacity, int kind) {
    return entries_bytes(capacity) + al(size_t(C) * N * record_bytes(kind));
}

// Largest bin table of the direct placement (its waves hold one table each in
// LDS, so larger tables cost occupancy; larger tile grids sort and pack).
constexpr size_t kPlaceLdsBytes = 16384;
constexpr uint32_t kPlaceWaves = 4096;  // target number of ranges (waves) per binning pass

static FusedScratch fused_scratch(uint32_t C, uint32_t N, uint32_t n_tiles, uint32_t capacity, int kind) {
    FusedScratch L = {};
    const size_t CN = size_t(C) * N;
    size_t o = 0;
    L.order_bounds = o; o += al(8 * sizeof(float));
    L.order_sort_n_bytes = sort_workspace_bytes(N);
    L.order_sort_n = o; o += al(L.order_sort_n_bytes);
    L.order_sorted_keys = o; o += al(size_t(N) * 4 + 4);
    L.params = o; o += al(size_t(N) * kPermutedFloats * sizeof(float));
    L.frames = o; o += al(size_t(N) * sizeof(PrimFrame));
    L.cams = o; o += al(size_t(C) * sizeof(CamFrame));
    L.rects = o; o += al(CN * sizeof(uint2));
    L.reach = o; o += al(CN * sizeof(float2));
    L.scan = o; o += al(scan_workspace_bytes(CN));
    const uint64_t nt = C ? n_tiles / C : 0;
    L.place = C > 0 && nt * kBinClasses * sizeof(uint32_t) <= kPlaceLdsBytes;
    if (L.place) {
        L.n_bins = uint32_t(nt) * kBinClasses;
        L.bin_bits = 32u - uint32_t(__builtin_clz(L.n_bins - 1u));
        const uint64_t per = (CN + kPlaceWaves - 1) / kPlaceWaves;
        L.range_len = uint32_t(std::max<uint64_t>(256, (per + 63) / 64 * 64));
        L.n_ranges = (N + L.range_len - 1) / L.range_len;
        L.range_counts = o; o += al(size_t(C) * L.n_ranges * L.n_bins * 4 + 4);
        L.bin_totals = o; o += al(size_t(C) * L.n_bins * 4 + 4);
        L.bin_scan = o; o += al(scan_workspace_bytes(uint64_t(C) * L.n_bins));
        L.bytes = o;
        return L;
    }
    L.sort_bytes = sort_workspace_bytes(capacity);
    L.sort = o; o += al(L.sort_bytes);
    L.keys = o; o += al(size_t(capacity) * 4 + 4);
    L.vals = o; o += al(size_t(capacity) * 4 + 4);
    L.tm = o; o += al(size_t(n_tiles) * 4 + 4);
    L.lengths = o; o += al(size_t(n_tiles) * 4 + 4);
    L.cm = o; o += al(size_t(n_tiles) * 4 + 8);
    L.tile_scan = o; o += al(scan_workspace_bytes(n_tiles));
    L.bytes = o;
    return L;
}

template <typename T>
static T* at(void* ws, size_t off) { return reinterpret_cast<T*>(static_cast<char*>(ws) + off); }

}  // namespace r2s

using namespace r2s;

extern "C" size_t r2s_fused_scratch_bytes(uint32_t C, uint32_t N, uint32_t width, uint32_t height,
                                          uint32_t tile_size, uint32_t capacity, int kind) {
    if (tile_size == 0) return 0;
    const uint64_t n_tiles = uint64_t(C) * ((width + tile_size - 1) / tile_size) *
                             ((height + tile_size - 1) / tile_size);
    if (n_tiles >= (1ull << 32)) return 0;
    return fused_scratch(C, N, uint32_t(n_tiles), capacity, kind).bytes;
}

extern "C" size_t r2s_fused_record_bytes(int kind) { return record_bytes(kind); }

extern "C" size_t r2s_fused_packed_bytes(uint32_t C, uint32_t N, uint32_t capacity, int kind) {
    return packed_bytes(C, N, capacity, kind);
}

static int fused_check(uint32_t C, uint32_t N, uint32_t ntx, uint32_t nty, uint32_t capacity) {
    R2S_REQUIRE(uint64_t(C) * N < (1ull << 31), "C*N exceeds the int32 index range");
    R2S_REQUIRE(uint64_t(C) * ntx * nty < (1ull << 32), "C*n_tiles exceeds the key range");
    R2S_REQUIRE(ntx < 65536 && nty < 65536, "tile grid too large");
    R2S_REQUIRE(capacity > 0 && capacity < (1u << 31), "capacity must be in (0, 2^31)");
    return R2S_OK;
}

static int check_tile(uint32_t tile_size, uint32_t waves) {
    R2S_REQUIRE(tile_size == 8 || tile_size == 16 || tile_size == 32, "fused tile_size must be 8, 16 or 32 (got %u)",
                tile_size);
    R2S_REQUIRE(waves == 0 || waves == 4 || waves == 8 || waves == 16, "waves per tile must be 4, 8 or 16 (got %u)",
                waves);
    return R2S_OK;
}

// Launches Launch<T, WAVES, BETA, VAR>::run for the runtime tile size, wave count and kind.
template <template <int, int, bool, bool> class Launch, typename... Args>
static void launch_tiled(uint32_t tile, uint32_t waves, int kind, Args&&... args) {
    if (waves == 0) waves = kDefaultWaves;
    const bool beta = kind & kKindBeta, var = kind & kKindVar;
#define R2S_TW(T_, W_)                                                           \
    if (tile == T_ && waves == W_) {                                             \
        if (beta && var) return Launch<T_, W_, true, true>::run(args...);        \
        if (beta) return Launch<T_, W_, true, false>::run(args...);              \
        if (var) return Launch<T_, W_, false, true>::run(args...);               \
        return Launch<T_, W_, false, false>::run(args...);                       \
    }
    R2S_TW(8, 4) R2S_TW(8, 8) R2S_TW(8, 16)
    R2S_TW(16, 4) R2S_TW(16, 8) R2S_TW(16, 16)
    R2S_TW(32, 4) R2S_TW(32, 8) R2S_TW(32, 16)
#undef R2S_TW
}

template <int T, int WAVES, bool BETA, bool VAR>
struct LaunchFwd {
    static void run(uint32_t n_tiles, hipStream_t stream, const void* packed, const int32_t* offsets,
                    const uint32_t* order, const uint32_t* counters, uint32_t capacity, uint32_t ntx, uint32_t nty,
                    uint32_t W, uint32_t H, float* image, float* image_var) {
        fused_raster_fwd_kernel<T, WAVES, BETA, VAR><<<n_tiles, WAVES * kWave, 0, stream>>>(
            static_cast<const uint2*>(packed),
            reinterpret_cast<const Stored<BETA, VAR>*>(static_cast<const char*>(packed) + entries_bytes(capacity)),
            offsets, order, counters, capacity, n_tiles, ntx, nty, W, H, image, image_var);
    }
};

template <int T, int WAVES, bool BETA, bool VAR>
struct LaunchBwd {
    static void run(uint32_t n_tiles, hipStream_t stream, const void* packed, const int32_t* offsets,
                    const uint32_t* order, const uint32_t* counters, uint32_t capacity, uint32_t ntx, uint32_t nty,
                    uint32_t W, uint32_t H, const float* v_image, const float* v_image_var, int detach,
                    float* isect_grads) {
        fused_raster_bwd_kernel<T, WAVES, BETA, VAR><<<n_tiles, WAVES * kWave, 0, stream>>>(
            static_cast<const uint2*>(packed),
            reinterpret_cast<const Stored<BETA, VAR>*>(static_cast<const char*>(packed) + entries_bytes(capacity)),
            offsets, order, counters, capacity, n_tiles, ntx, nty, W, H, v_image, v_image_var, detach, isect_grads);
    }
};

template <int T, int WAVES, bool BETA, bool VAR>
struct LaunchFwdPacked {
    static void run(uint32_t n_groups, hipStream_t stream, const r2s_raster_job* jobs, uint32_t n_jobs,
                    uint32_t total) {
        fused_raster_fwd_packed_kernel<T, WAVES, BETA, VAR><<<n_groups, WAVES * kWave, 0, stream>>>(jobs, n_jobs,
                                                                                                      total);
    }
};

template <int T, int WAVES, bool BETA, bool VAR>
struct LaunchBwdPacked {
    static void run(uint32_t n_groups, hipStream_t stream, const r2s_raster_job* jobs, uint32_t n_jobs,
                    uint32_t total) {
        fused_raster_bwd_packed_kernel<T, WAVES, BETA, VAR><<<n_groups, WAVES * kWave, 0, stream>>>(jobs, n_jobs,
                                                                                                      total);
    }
};

template <int T, int WAVES, bool BETA, bool VAR>
struct LaunchKeep {
    static void run(uint32_t n_tiles, hipStream_t stream, const void* packed, const int32_t* offsets,
                    const uint32_t* counters, uint32_t capacity, uint32_t ntx, uint32_t nty, uint32_t* keep_fwd,
                    uint32_t* keep_bwd) {
        fused_keep_mask_kernel<T, BETA, VAR><<<n_tiles, 256, 0, stream>>>(
            static_cast<const uint2*>(packed),
            reinterpret_cast<const Stored<BETA, VAR>*>(static_cast<const char*>(packed) + entries_bytes(capacity)),
            offsets, counters, capacity, n_tiles, ntx, nty, keep_fwd, keep_bwd);
    }
};

template <bool BETA, bool VAR>
static int prepare_impl(const float* means, const float* quats, const float* scales, const float* densities,
                        const float* betas, const float* beta_ratio, const float* uncertainties,
                        const float* viewmats, const float* Ks, uint32_t C, uint32_t N, uint32_t width,
                        uint32_t height, uint32_t tile_size, float eps, float radius_clip, uint32_t capacity,
                        void* scratch, size_t scratch_bytes, void* packed, int32_t* tile_offsets,
                        uint32_t* tile_order, uint32_t* record_offsets, uint32_t* counters, hipStream_t stream) {
    using R = Stored<BETA, VAR>;
    const uint32_t ntx = (width + tile_size - 1) / tile_size, nty = (height + tile_size - 1) / tile_size;
    if (int rc = fused_check(C, N, ntx, nty, capacity)) return rc;
    const FusedScratch L = fused_scratch(C, N, C * ntx * nty, capacity, (BETA ? kKindBeta : 0) | (VAR ? kKindVar : 0));
    R2S_REQUIRE(scratch_bytes >= L.bytes, "fused scratch too small: %zu < %zu", scratch_bytes, L.bytes);
    const uint32_t CN = C * N, n_tiles = C * ntx * nty;
    if (int rc = fill_zero<uint32_t>(counters, 2, stream)) return rc;
    uint2* entries = static_cast<uint2*>(packed);
    R* recs = at<R>(packed, entries_bytes(capacity));
    uint32_t* cum = record_offsets;
    uint32_t* perm = record_offsets + size_t(CN) + 1;  // the primitive order (input index per position)
    float* pp = at<float>(scratch, L.params);
    float *p_means = pp, *p_dens = pp + 3 * size_t(N), *p_betas = pp + 4 * size_t(N), *p_ratio = pp + 5 * size_t(N);
    float* p_unc = pp + 6 * size_t(N);
    PrimFrame* p_frames = at<PrimFrame>(scratch, L.frames);
    if (N > 0) {
        void* ows = at<void>(scratch, L.order_sort_n);
        uint32_t* bounds = at<uint32_t>(scratch, L.order_bounds);
        morton_bounds_init_kernel<<<1, kWave, 0, stream>>>(bounds);
        if (int rc = check_launch("morton_bounds_init")) return rc;
        morton_bounds_kernel<<<std::min(ceil_div(N, kOrderThreads), 256u), kOrderThreads, 0, stream>>>(means, N,
                                                                                                       bounds);
        if (int rc = check_launch("morton_bounds")) return rc;
        // small sets sort by the code's top 8 bits only (one radix pass; 8 x 8 x 4
        // cells still keep a range of neighbouring primitives on a few tiles),
        // larger ones by all 15 (two passes; coarser cells cost the raster
        // kernels' gather locality there)
        const uint32_t code_bits = N <= kMortonCoarseMax ? 8u : 3u * kMortonAxisBits;
        morton_keys_kernel<<<ceil_div(N, 256), 256, 0, stream>>>(means, N, bounds, 3u * kMortonAxisBits - code_bits,
                                                                 sort_keys_buffer(ows), sort_vals_buffer(ows, N));
        if (int rc = check_launch("morton_keys")) return rc;
        if (int rc = sort_pairs_u32(ows, L.order_sort_n_bytes, N, nullptr, code_bits,
                                    at<uint32_t>(scratch, L.order_sorted_keys), perm, stream))
            return rc;
        permute_params_kernel<<<ceil_div(N, 256), 256, 0, stream>>>(
            perm, N, means, quats, scales, densities, BETA ? betas : nullptr, BETA ? beta_ratio : nullptr,
            VAR ? uncertainties : nullptr, p_means, p_dens, p_betas, p_ratio, p_unc, p_frames, viewmats, Ks, C,
            width, height, at<CamFrame>(scratch, L.cams));
        if (int rc = check_launch("permute_params")) return rc;
    }
    if (CN > 0) {  // (N > 0: the permute pass wrote the camera frames)
        const CamFrame* cams = at<CamFrame>(scratch, L.cams);
        dim3 grid(ceil_div(N, 256), ceil_div(C, kProjectCams));
        fused_project_kernel<BETA, VAR><<<grid, 256, 0, stream>>>(
            p_means, p_frames, p_dens, p_betas, p_ratio, p_unc, cams, C, width, height, eps, radius_clip,
            tile_size, ntx, nty, N, recs, at<uint2>(scratch, L.rects), at<float2>(scratch, L.reach), cum);
        if (int rc = check_launch("fused_project")) return rc;
    }
    if (int rc = scan_exclusive(cum, CN, cum, counters, at<void>(scratch, L.scan), stream)) return rc;
    if (L.place) {
        // (with CN = 0 the count is 0 and the zeroed overflow flag stays 0)
        uint32_t* range_counts = at<uint32_t>(scratch, L.range_counts);
        uint32_t* totals = at<uint32_t>(scratch, L.bin_totals);
        const size_t lds = size_t(L.n_bins) * sizeof(uint32_t);
        const dim3 grid(L.n_ranges, C);
        if (CN > 0) {
            fused_bin_count_kernel<<<grid, kWave, lds, stream>>>(at<uint2>(scratch, L.rects),
                                                                 at<float2>(scratch, L.reach), N, L.range_len,
                                                                 L.n_ranges, ntx, tile_size, L.n_bins, range_counts,
                                                                 counters, capacity);
            if (int rc = check_launch("fused_bin_count")) return rc;
        }
        const uint32_t n_cols = C * L.n_bins;
        fused_bin_scan_ranges_kernel<<<ceil_div(n_cols, 256), 256, 0, stream>>>(range_counts, C, L.n_ranges, L.n_bins,
                                                                               totals);
        if (int rc = check_launch("fused_bin_scan_ranges")) return rc;
        if (int rc = scan_exclusive(totals, n_cols, totals, nullptr, at<void>(scratch, L.bin_scan), stream)) return rc;
        if (CN > 0) {
            fused_bin_place_kernel<<<grid, kWave, lds, stream>>>(
                at<uint2>(scratch, L.rects), at<float2>(scratch, L.reach), cum, N,
                L.range_len, L.n_ranges, ntx, tile_size, L.n_bins, L.bin_bits, range_counts, totals, capacity,
                entries);
            if (int rc = check_launch("fused_bin_place")) return rc;
        }
        if (tile_order && n_tiles > 0) {  // the tile offsets come with the tile order
            fused_tile_order_kernel<<<8, kTileOrderThreads, 0, stream>>>(totals, kBinClasses, counters, capacity,
                                                                         n_tiles, tile_order, tile_offsets);
            return check_launch("fused_tile_order");
        }
        fused_bin_tile_offsets_kernel<<<ceil_div(n_tiles, 256), 256, 0, stream>>>(totals, n_tiles, tile_offsets);
        if (int rc = check_launch("fused_bin_tile_offsets")) return rc;
    } else {
    void* sort_ws = at<void>(scratch, L.sort);
    if (CN > 0) {
        fused_emit_kernel<<<ceil_div(CN, 256), 256, 0, stream>>>(
            at<uint2>(scratch, L.rects), cum, CN, ntx, capacity, sort_keys_buffer(sort_ws),
            sort_vals_buffer(sort_ws, capacity), counters + 1);
        if (int rc = check_launch("fused_emit")) return rc;
    }
    uint32_t* keys = at<uint32_t>(scratch, L.keys);
    uint32_t* vals = at<uint32_t>(scratch, L.vals);
    if (int rc = sort_pairs_u32(sort_ws, L.sort_bytes, capacity, counters, key_end_bit(ntx * nty - 1), keys, vals,
                                stream))
        return rc;
    uint32_t* tm = at<uint32_t>(scratch, L.tm);
    uint32_t* cm = at<uint32_t>(scratch, L.cm);
    fused_offsets_kernel<<<ceil_div(n_tiles, 256), 256, 0, stream>>>(keys, vals, counters, capacity, C, N, n_tiles,
                                                                     reinterpret_cast<int32_t*>(tm));
    if (int rc = check_launch("fused_offsets")) return rc;
    fused_list_lengths_kernel<<<ceil_div(n_tiles, 256), 256, 0, stream>>>(tm, counters, capacity, C, ntx * nty,
                                                                          at<uint32_t>(scratch, L.lengths));
    if (int rc = check_launch("fused_list_lengths")) return rc;
    if (int rc = scan_exclusive(at<uint32_t>(scratch, L.lengths), n_tiles, cm, nullptr, at<void>(scratch, L.tile_scan),
                                stream))
        return rc;
    const uint32_t pack_blocks = (n_tiles + 7u) & ~7u;  // a multiple of the 8 XCDs
    fused_pack_kernel<<<pack_blocks, kPackThreads, 0, stream>>>(
        at<uint2>(scratch, L.rects), at<float2>(scratch, L.reach), cum, CN, C, ntx * nty, ntx, tile_size, vals, tm,
        cm, counters, capacity, entries, tile_offsets);
    if (int rc = check_launch("fused_pack")) return rc;
    }
    if (tile_order && n_tiles > 0) {
        fused_tile_order_kernel<<<8, kTileOrderThreads, 0, stream>>>(reinterpret_cast<const uint32_t*>(tile_offsets),
                                                                     1u, counters, capacity, n_tiles, tile_order,
                                                                     nullptr);
        if (int rc = check_launch("fused_tile_order")) return rc;
    }
    return R2S_OK;
}

extern "C" int r2s_fused_prepare(const float* means, const float* quats, const float* scales,
                                 const float* densities, const float* betas, const float* beta_ratio,
                                 const float* uncertainties, const float* viewmats, const float* Ks, uint32_t C,
                                 uint32_t N, uint32_t width, uint32_t height, uint32_t tile_size, float eps,
                                 float radius_clip, uint32_t capacity, void* scratch, size_t scratch_bytes,
                                 void* packed, int32_t* tile_offsets, uint32_t* tile_order, uint32_t* record_offsets,
                                 uint32_t* counters, r2s_stream_t stream) {
    if (int rc = check_tile(tile_size, 0)) return rc;
    R2S_REQUIRE((betas == nullptr) == (beta_ratio == nullptr), "betas and beta_ratio go together");
#define R2S_PREP(B_, V_)                                                                                         \
    return prepare_impl<B_, V_>(means, quats, scales, densities, betas, beta_ratio, uncertainties, viewmats, Ks, C, \
                                N, width, height, tile_size, eps, radius_clip, capacity, scratch, scratch_bytes,     \
                                packed, tile_offsets, tile_order, record_offsets, counters, stream)
    if (betas && uncertainties) R2S_PREP(true, true);
    if (betas) R2S_PREP(true, false);
    if (uncertainties) R2S_PREP(false, true);
    R2S_PREP(false, false);
#undef R2S_PREP
}

extern "C" int r2s_fused_render(const void* packed, const int32_t* tile_offsets, const uint32_t* tile_order,
                                const uint32_t* counters, uint32_t capacity, uint32_t C, uint32_t width,
                                uint32_t height, uint32_t tile_size, uint32_t waves, int kind, float* image,
                                float* image_var, r2s_stream_t stream) {
    if (int rc = check_tile(tile_size, waves)) return rc;
    R2S_REQUIRE(!(kind & kKindVar) || image_var, "a variance kind needs image_var");
    const uint32_t ntx = (width + tile_size - 1) / tile_size, nty = (height + tile_size - 1) / tile_size;
    const uint64_t n_tiles = uint64_t(C) * ntx * nty;
    R2S_REQUIRE(n_tiles < (1ull << 31), "too many tiles for one launch");
    if (n_tiles == 0) return R2S_OK;
    launch_tiled<LaunchFwd>(tile_size, waves, kind, uint32_t(n_tiles), stream, packed, tile_offsets, tile_order,
                            counters, capacity, ntx, nty, width, height, image, image_var);
    return check_launch("fused_raster_fwd");
}

extern "C" int r2s_fused_raster_vjp(const void* packed, const int32_t* tile_offsets, const uint32_t* tile_order,
                                    const uint32_t* counters, uint32_t capacity, uint32_t C, uint32_t width,
                                    uint32_t height, uint32_t tile_size, uint32_t waves, int kind,
                                    const float* v_image, const float* v_image_var, int detach, float* isect_grads,
                                    r2s_stream_t stream) {
    if (int rc = check_tile(tile_size, waves)) return rc;
    const uint32_t ntx = (width + tile_size - 1) / tile_size, nty = (height + tile_size - 1) / tile_size;
    const uint64_t n_tiles = uint64_t(C) * ntx * nty;
    R2S_REQUIRE(n_tiles < (1ull << 31), "too many tiles for one launch");
    if (n_tiles == 0) return R2S_OK;
    launch_tiled<LaunchBwd>(tile_size, waves, kind, uint32_t(n_tiles), stream, packed, tile_offsets, tile_order,
                            counters, capacity, ntx, nty, width, height, v_image, v_image_var, detach, isect_grads);
    return check_launch("fused_raster_bwd");
}

// Checks a packed job table (host copy); returns the total tile count and the
// grid (8 x the longest per-XCD sequence of job bands).
static int check_packed(const r2s_raster_job* jobs, const r2s_raster_job* jobs_device, uint32_t n_jobs,
                        uint32_t tile_size, uint32_t waves, int kind, bool vjp, uint32_t* total,
                        uint32_t* groups) {
    if (int rc = check_tile(tile_size, waves)) return rc;
    R2S_REQUIRE(jobs && jobs_device, "packed: both job tables are required");
    R2S_REQUIRE(n_jobs >= 1 && n_jobs <= 1024, "packed: 1 to 1024 jobs (got %u)", n_jobs);
    uint64_t t = 0, seq[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t j = 0; j < n_jobs; ++j) {
        const r2s_raster_job& J = jobs[j];
        R2S_REQUIRE(J.first_tile == t, "packed: job %u starts at tile %u, expected %llu", j, J.first_tile,
                    (unsigned long long)t);
        R2S_REQUIRE(J.packed && J.tile_offsets && J.counters && J.capacity > 0, "packed: job %u lacks its buffers", j);
        if (vjp)
            R2S_REQUIRE(J.v_image && J.isect_grads, "packed: job %u lacks v_image or isect_grads", j);
        else
            R2S_REQUIRE(J.image && (!(kind & kKindVar) || J.image_var), "packed: job %u lacks its image", j);
        const uint64_t ntx = (J.width + tile_size - 1) / tile_size, nty = (J.height + tile_size - 1) / tile_size;
        const uint64_t n = uint64_t(J.C) * ntx * nty;
        for (uint32_t x = 0; x < 8; ++x) seq[x] += n / 8 + (x < n % 8 ? 1 : 0);
        t += n;
    }
    R2S_REQUIRE(t + 8ull * n_jobs < (1ull << 31), "packed: too many tiles for one launch");
    *total = uint32_t(t);
    *groups = 8u * uint32_t(*std::max_element(seq, seq + 8));
    return R2S_OK;
}

extern "C" int r2s_fused_render_packed(const r2s_raster_job* jobs_host, const r2s_raster_job* jobs_device,
                                       uint32_t n_jobs, uint32_t tile_size, uint32_t waves, int kind,
                                       r2s_stream_t stream) {
    uint32_t total = 0, groups = 0;
    if (int rc = check_packed(jobs_host, jobs_device, n_jobs, tile_size, waves, kind, false, &total, &groups))
        return rc;
    if (total == 0) return R2S_OK;
    launch_tiled<LaunchFwdPacked>(tile_size, waves, kind, groups, stream, jobs_device, n_jobs, total);
    return check_launch("fused_raster_fwd_packed");
}

extern "C" int r2s_fused_raster_vjp_packed(const r2s_raster_job* jobs_host, const r2s_raster_job* jobs_device,
                                           uint32_t n_jobs, uint32_t tile_size, uint32_t waves, int kind,
                                           r2s_stream_t stream) {
    uint32_t total = 0, groups = 0;
    if (int rc = check_packed(jobs_host, jobs_device, n_jobs, tile_size, waves, kind, true, &total, &groups))
        return rc;
    if (total == 0) return R2S_OK;
    launch_tiled<LaunchBwdPacked>(tile_size, waves, kind, groups, stream, jobs_device, n_jobs, total);
    return check_launch("fused_raster_bwd_packed");
}

extern "C" int r2s_fused_keep_masks(const void* packed, const int32_t* tile_offsets, const uint32_t* counters,
                                    uint32_t capacity, uint32_t C, uint32_t width, uint32_t height,
                                    uint32_t tile_size, int kind, uint32_t* keep_fwd, uint32_t* keep_bwd,
                                    r2s_stream_t stream) {
    if (int rc = check_tile(tile_size, 0)) return rc;
    R2S_REQUIRE(keep_fwd && keep_bwd, "keep_masks: both mask buffers are required");
    const uint32_t ntx = (width + tile_size - 1) / tile_size, nty = (height + tile_size - 1) / tile_size;
    const uint64_t n_tiles = uint64_t(C) * ntx * nty;
    R2S_REQUIRE(n_tiles < (1ull << 31), "too many tiles for one launch");
    if (n_tiles == 0) return R2S_OK;
    // the wave count is not part of the decisions; the 8-wave instantiation is a placeholder
    launch_tiled<LaunchKeep>(tile_size, 8, kind, uint32_t(n_tiles), stream, packed, tile_offsets, counters,
                             capacity, ntx, nty, keep_fwd, keep_bwd);
    return check_launch("fused_keep_masks");
}

// The one-pass parameter VJP needs no workspace; the size query and the
// workspace arguments stay for callers that pass one.
extern "C" size_t r2s_fused_param_workspace_bytes(uint32_t, uint32_t) { return 0; }

extern "C" int r2s_fused_param_vjp(const float* means, const float* quats, const float* scales,
                                   const float* densities, const float* betas, const float* beta_ratio,
                                   const float* beta_log_grad, const float* viewmats, const float* Ks, uint32_t C,
                                   uint32_t N, uint32_t width, uint32_t height, float eps,
                                   const uint32_t* record_offsets, const uint32_t* counters, uint32_t capacity,
                                   const float* isect_grads, void* workspace, size_t workspace_bytes,
                                   float* v_means, float* v_quats, float* v_scales, float* v_densities,
                                   float* v_betas, float* v_uncertainties, r2s_stream_t stream) {
    (void)workspace;
    (void)workspace_bytes;
    if (N == 0) return R2S_OK;
    if (int rc = fused_check(C, N, 1, 1, capacity)) return rc;
    const bool beta = betas != nullptr;
    R2S_REQUIRE(!beta || (beta_ratio && beta_log_grad && v_betas), "Beta primitives need ratio, log-gradient and "
                "v_betas buffers");
    const uint32_t blocks = ceil_div(N, uint32_t(kWave));
#define R2S_VJP(B_, V_)                                                                                          \
    fused_param_vjp_kernel<B_, V_><<<blocks, kVjpWaves * kWave, 0, stream>>>(                                   \
        means, quats, scales, densities, B_ ? beta_ratio : nullptr, B_ ? beta_log_grad : nullptr, viewmats, Ks,  \
        width, height, eps, C, N, record_offsets, counters, capacity, isect_grads, v_means, v_quats, v_scales,  \
        v_densities, B_ ? v_betas : nullptr, v_uncertainties, record_offsets + size_t(C) * N + 1)
    if (beta && v_uncertainties) R2S_VJP(true, true);
    else if (beta) R2S_VJP(true, false);
    else if (v_uncertainties) R2S_VJP(false, true);
    else R2S_VJP(false, false);
#undef R2S_VJP
    return check_launch("fused_param_vjp");
}

extern "C" int r2s_fused_forward(const float* means, const float* quats, const float* scales,
                                 const float* densities, const float* betas, const float* beta_ratio,
                                 const float* uncertainties, const float* viewmats, const float* Ks, uint32_t C,
                                 uint32_t N, uint32_t width, uint32_t height, uint32_t tile_size, float eps,
                                 float radius_clip, uint32_t capacity, void* scratch, size_t scratch_bytes,
                                 void* packed, int32_t* tile_offsets, uint32_t* tile_order, uint32_t* record_offsets,
                                 uint32_t* counters, float* image, float* image_var, r2s_stream_t stream) {
    if (int rc = r2s_fused_prepare(means, quats, scales, densities, betas, beta_ratio, uncertainties, viewmats, Ks, C,
                                   N, width, height, tile_size, eps, radius_clip, capacity, scratch, scratch_bytes,
                                   packed, tile_offsets, tile_order, record_offsets, counters, stream))
        return rc;
    const int kind = (betas ? kKindBeta : 0) | (uncertainties ? kKindVar : 0);
    return r2s_fused_render(packed, tile_offsets, tile_order, counters, capacity, C, width, height, tile_size, 0,
                            kind, image, image_var, stream);
}

extern "C" int r2s_fused_backward(const float* means, const float* quats, const float* scales,
                                  const float* densities, const float* betas, const float* beta_ratio,
                                  const float* beta_log_grad, const float* uncertainties, const float* viewmats,
                                  const float* Ks, uint32_t C, uint32_t N, uint32_t width, uint32_t height,
                                  uint32_t tile_size, float eps, uint32_t capacity, const void* packed,
                                  const int32_t* tile_offsets, const uint32_t* tile_order,
                                  const uint32_t* record_offsets, const uint32_t* counters, float* isect_grads,
                                  void* workspace, size_t workspace_bytes, const float* v_image,
                                  const float* v_image_var, int detach, float* v_means, float* v_quats,
                                  float* v_scales, float* v_densities, float* v_betas, float* v_uncertainties,
                                  r2s_stream_t stream) {
    const int kind = (betas ? kKindBeta : 0) | (uncertainties ? kKindVar : 0);
    if (int rc = r2s_fused_raster_vjp(packed, tile_offsets, tile_order, counters, capacity, C, width, height,
                                      tile_size, 0, kind, v_image, v_image_var, detach, isect_grads, stream))
        return rc;
    return r2s_fused_param_vjp(means, quats, scales, densities, betas, beta_ratio, beta_log_grad, viewmats, Ks, C, N,
                               width, height, eps, record_offsets, counters, capacity, isect_grads, workspace,
                               workspace_bytes, v_means, v_quats, v_scales, v_densities, v_betas,
                               uncertainties ? v_uncertainties : nullptr, stream);
}

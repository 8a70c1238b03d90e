// Tile binning: per-primitive tile rectangles, a device-wide exclusive scan,
// record emission, a stable LSD radix sort and per-tile offsets.
//
// The sort is written for wave64: each 8-bit digit pass ranks keys inside a
// wave with ballot "multisplit" (eight ballots give the mask of lanes holding
// the same digit), keeps per-wave running digit counters in LDS, stages the
// block's keys in LDS in sorted order and scatters them in digit runs. Ranks
// follow input order, so equal keys keep their relative order (stable), which
// makes the per-tile primitive order ascending in cam*N + g as in the
// reference's CUB sort (r2splat/cuda/csrc/IsectTilesKernel.cu:398-445).
#include "common.hpp"
#include "binning.hpp"

namespace r2s {

// ---------------------------------------------------------------------------
// Tile rectangle of one primitive (IsectTilesKernel.cu:30-40,91-112 and
// IsectTiles3DKernel.cu:36-45,~95-110): bounds are floor/ceil of
// (mean -/+ radius) / tile_size, clamped to [0, n_tiles] before conversion.
// ---------------------------------------------------------------------------
template <typename S>
__device__ __forceinline__ bool rect_2d(const S* __restrict__ means2d, const int32_t* __restrict__ radii,
                                        bool rect, size_t idx, uint32_t tile, uint32_t ntx,
                                        uint32_t nty, uint32_t& x0, uint32_t& x1, uint32_t& y0,
                                        uint32_t& y1) {
    const int32_t rx = rect ? radii[2 * idx] : radii[idx];
    const int32_t ry = rect ? radii[2 * idx + 1] : radii[idx];
    if (rx <= 0 || ry <= 0) return false;
    const S mx = means2d[2 * idx], my = means2d[2 * idx + 1];
    if (!isfinite(mx) || !isfinite(my)) return false;
    const bool okx = tile_span(mx, rx, tile, ntx, x0, x1);
    const bool oky = tile_span(my, ry, tile, nty, y0, y1);
    return okx && oky;
}

template <typename S>
__global__ void __launch_bounds__(256)
count_2d_kernel(const S* __restrict__ means2d, const int32_t* __restrict__ radii, bool rect,
                uint32_t total, uint32_t tile, uint32_t ntx, uint32_t nty,
                int32_t* __restrict__ tiles_per_gauss, uint32_t* __restrict__ counts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    uint32_t x0, x1, y0, y1, cnt = 0;
    if (rect_2d(means2d, radii, rect, i, tile, ntx, nty, x0, x1, y0, y1)) cnt = (x1 - x0) * (y1 - y0);
    tiles_per_gauss[i] = int32_t(cnt);
    counts[i] = cnt;
}

template <typename S>
__global__ void __launch_bounds__(256)
emit_2d_kernel(const S* __restrict__ means2d, const int32_t* __restrict__ radii, bool rect,
               uint32_t total, uint32_t N, uint32_t tile, uint32_t ntx, uint32_t nty,
               const uint32_t* __restrict__ cum, uint32_t* __restrict__ keys,
               uint32_t* __restrict__ vals) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    uint32_t x0, x1, y0, y1;
    if (!rect_2d(means2d, radii, rect, i, tile, ntx, nty, x0, x1, y0, y1)) return;
    const uint32_t cam_base = (i / N) * (ntx * nty);
    uint32_t w = cum[i];
    for (uint32_t y = y0; y < y1; ++y)
        for (uint32_t x = x0; x < x1; ++x, ++w) {
            keys[w] = cam_base + y * ntx + x;
            vals[w] = i;
        }
}

template <typename S>
__global__ void __launch_bounds__(256)
count_3d_kernel(const S* __restrict__ means3d, const int32_t* __restrict__ radii, uint32_t N,
                uint32_t tile, uint32_t ntx, uint32_t nty, uint32_t ntz,
                int32_t* __restrict__ tiles_per_gauss, uint32_t* __restrict__ counts) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    uint32_t lo[3], hi[3], cnt = 0;
    if (box_3d(means3d, radii, n, tile, ntx, nty, ntz, lo, hi))
        cnt = (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
    tiles_per_gauss[n] = int32_t(cnt);
    counts[n] = cnt;
}

template <typename S>
__global__ void __launch_bounds__(256)
emit_3d_kernel(const S* __restrict__ means3d, const int32_t* __restrict__ radii, uint32_t N,
               uint32_t tile, uint32_t ntx, uint32_t nty, uint32_t ntz,
               const uint32_t* __restrict__ cum, uint32_t* __restrict__ keys,
               uint32_t* __restrict__ vals) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    uint32_t lo[3], hi[3];
    if (!box_3d(means3d, radii, n, tile, ntx, nty, ntz, lo, hi)) return;
    uint32_t w = cum[n];
    for (uint32_t z = lo[2]; z < hi[2]; ++z)
        for (uint32_t y = lo[1]; y < hi[1]; ++y)
            for (uint32_t x = lo[0]; x < hi[0]; ++x, ++w) {
                keys[w] = (z * nty + y) * ntx + x;
                vals[w] = n;
            }
}

// ---------------------------------------------------------------------------
// Device-wide exclusive scan of uint32 (reduce -> scan partials -> apply).
// out[0..n] with out[n] = total; out may alias in.
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr uint32_t kScanTile = kScanThreads * kScanItems;
// arrays up to 16384 elements (and the partials of arrays up to 2^25) are
// scanned by one workgroup of 1024 threads in a single launch
constexpr int kScanBigThreads = 1024;
constexpr int kScanBigItems = 16;
constexpr uint32_t kScanBigTile = kScanBigThreads * kScanBigItems;

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x) {
    const int lane = lane_id();
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, kWave);
        if (lane >= off) x += y;
    }
    return x;
}

// Block-wide exclusive scan of one value per thread; returns the block total.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t x, uint32_t& excl, uint32_t* lds_waves) {
    const int lane = lane_id(), wave = threadIdx.x / kWave;
    const int n_waves = blockDim.x / kWave;
    const uint32_t incl = wave_inclusive_scan(x);
    if (lane == kWave - 1) lds_waves[wave] = incl;
    __syncthreads();
    uint32_t base = 0, total = 0;
    for (int w = 0; w < n_waves; ++w) {
        const uint32_t v = lds_waves[w];
        if (w < wave) base += v;
        total += v;
    }
    __syncthreads();
    excl = base + incl - x;
    return total;
}

__global__ void __launch_bounds__(kScanThreads)
scan_reduce_kernel(const uint32_t* __restrict__ in, uint32_t n, uint32_t* __restrict__ partial) {
    __shared__ uint32_t lds[kScanThreads / kWave];
    const size_t base = size_t(blockIdx.x) * kScanTile;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const size_t i = base + size_t(k) * kScanThreads + threadIdx.x;
        if (i < n) s += in[i];
    }
    s = wave_sum(s);
    if (lane_id() == 0) lds[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kScanThreads / kWave; ++w) t += lds[w];
        partial[blockIdx.x] = t;
    }
}

// Scans one tile of THREADS x ITEMS elements per block; block b adds base[b]
// (if given). The block holding the last element writes out[n] = total (and
// *d_total when given).
template <int THREADS, int ITEMS>
__global__ void __launch_bounds__(THREADS)
scan_tile_kernel(const uint32_t* __restrict__ in, uint32_t n, const uint32_t* __restrict__ base,
                 uint32_t* __restrict__ out, uint32_t* __restrict__ d_total) {
    constexpr uint32_t kTile = uint32_t(THREADS) * ITEMS;
    __shared__ uint32_t lds[THREADS / kWave];
    const size_t tile0 = size_t(blockIdx.x) * kTile;
    const size_t t0 = tile0 + size_t(threadIdx.x) * ITEMS;
    uint32_t v[ITEMS], s = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        v[k] = (t0 + k < n) ? in[t0 + k] : 0u;
        s += v[k];
    }
    uint32_t excl;
    const uint32_t total = block_exclusive_scan(s, excl, lds);
    const uint32_t b = base ? base[blockIdx.x] : 0u;
    uint32_t run = b + excl;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        if (t0 + k < n) out[t0 + k] = run;
        run += v[k];
    }
    const bool last_block = (tile0 + kTile >= n);
    if (last_block && threadIdx.x == 0) {
        out[n] = b + total;
        if (d_total) *d_total = b + total;
    }
}

static size_t scan_levels_bytes(uint64_t n) {
    size_t bytes = 0;
    while (n > kScanTile) {
        n = (n + kScanTile - 1) / kScanTile;
        bytes += (n + 1) * sizeof(uint32_t);
        bytes = (bytes + 255) & ~size_t(255);
    }
    return bytes + 256;
}

int scan_exclusive(const uint32_t* in, uint32_t n, uint32_t* out, uint32_t* d_total, void* ws,
                   hipStream_t stream) {
    if (n <= kScanTile) {
        scan_tile_kernel<kScanThreads, kScanItems><<<1, kScanThreads, 0, stream>>>(in, n, nullptr, out, d_total);
        return check_launch("scan_tile");
    }
    if (n <= kScanBigTile) {
        scan_tile_kernel<kScanBigThreads, kScanBigItems><<<1, kScanBigThreads, 0, stream>>>(in, n, nullptr, out,
                                                                                            d_total);
        return check_launch("scan_block");
    }
    const uint32_t nb = ceil_div(n, kScanTile);
    uint32_t* partial = static_cast<uint32_t*>(ws);
    size_t used = ((size_t(nb) + 1) * sizeof(uint32_t) + 255) & ~size_t(255);
    scan_reduce_kernel<<<nb, kScanThreads, 0, stream>>>(in, n, partial);
    if (int rc = check_launch("scan_reduce")) return rc;
    if (int rc = scan_exclusive(partial, nb, partial, nullptr, static_cast<char*>(ws) + used, stream)) return rc;
    scan_tile_kernel<kScanThreads, kScanItems><<<nb, kScanThreads, 0, stream>>>(in, n, partial, out, d_total);
    return check_launch("scan_apply");
}

size_t scan_workspace_bytes(uint64_t n) { return scan_levels_bytes(n); }

// ---------------------------------------------------------------------------
// Stable LSD radix sort of (uint32 key, uint32 value) pairs.
// ---------------------------------------------------------------------------
constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / kWave;
constexpr int kSortRounds = 16;  // rounds of 64 keys per wave
constexpr uint32_t kSortTile = kSortThreads * kSortRounds;  // 4096 keys per block
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;

// `n` is the element count, or the capacity when `d_n` (device count) is given.
// A device count above the capacity means the emitted records are incomplete:
// the sort then treats the input as empty (the caller re-runs with more room).
__device__ __forceinline__ uint32_t live_count(uint32_t n, const uint32_t* d_n) {
    return d_n ? ((*d_n <= n) ? *d_n : 0u) : n;
}

// Per-block digit histogram. Lanes of a wave holding the same digit are
// grouped with ballots and counted once by their lowest lane, so runs of equal
// keys (common in emission order) cost one LDS atomic per run, not per key.
__global__ void __launch_bounds__(kSortThreads)
radix_upsweep_kernel(const uint32_t* __restrict__ keys, uint32_t n, const uint32_t* __restrict__ d_n,
                     uint32_t shift, uint32_t nbits, uint32_t n_blocks, uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[kRadix];
    n = live_count(n, d_n);
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t mask = (1u << nbits) - 1u;
    const int lane = lane_id();
    const size_t base = size_t(blockIdx.x) * kSortTile;
    for (int k = 0; k < kSortRounds; ++k) {
        const size_t i = base + size_t(k) * kSortThreads + threadIdx.x;
        const bool valid = i < n;
        const uint32_t d = valid ? (keys[i] >> shift) & mask : 0u;
        uint64_t peers = __ballot(valid);
        if (!peers) continue;  // wave-uniform
        for (uint32_t b = 0; b < nbits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const int leader = __ffsll((unsigned long long)peers) - 1;
        if (valid && lane == leader) atomicAdd(&h[d], uint32_t(__popcll(peers)));
    }
    __syncthreads();
    hist[size_t(threadIdx.x) * n_blocks + blockIdx.x] = h[threadIdx.x];
}

template <typename KOut>
__global__ void __launch_bounds__(kSortThreads)
radix_downsweep_kernel(const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
                       KOut* __restrict__ keys_out, uint32_t* __restrict__ vals_out, uint32_t n,
                       const uint32_t* __restrict__ d_n, uint32_t shift, uint32_t nbits,
                       uint32_t n_blocks, const uint32_t* __restrict__ hist_scanned) {
    n = live_count(n, d_n);
    if (size_t(blockIdx.x) * kSortTile >= n) return;  // whole block beyond the live count
    __shared__ uint32_t wave_cnt[kSortWaves][kRadix];
    __shared__ uint32_t local_off[kRadix];
    __shared__ uint32_t global_base[kRadix];
    __shared__ uint32_t scan_tmp[kSortWaves];
    __shared__ uint32_t s_keys[kSortTile];
    __shared__ uint32_t s_vals[kSortTile];

    const int lane = lane_id(), wave = threadIdx.x / kWave;
    const uint32_t mask = (1u << nbits) - 1u;
    for (int w = 0; w < kSortWaves; ++w) wave_cnt[w][threadIdx.x] = 0;
    __syncthreads();

    const size_t tile0 = size_t(blockIdx.x) * kSortTile;
    const size_t wave0 = tile0 + size_t(wave) * (kSortRounds * kWave);
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t key[kSortRounds], val[kSortRounds], rank[kSortRounds];
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r) {
        const size_t i = wave0 + size_t(r) * kWave + lane;
        const bool valid = i < n;
        key[r] = valid ? keys_in[i] : 0u;
        val[r] = valid ? vals_in[i] : 0u;
        const uint32_t d = (key[r] >> shift) & mask;
        uint64_t peers = __ballot(valid);
        for (uint32_t b = 0; b < nbits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        uint32_t old = 0;
        const int leader = peers ? (__ffsll((unsigned long long)peers) - 1) : 0;
        if (valid && lane == leader) old = atomicAdd(&wave_cnt[wave][d], uint32_t(__popcll(peers)));
        old = __shfl(old, leader, kWave);
        rank[r] = old + uint32_t(__popcll(peers & lt_mask));
    }
    __syncthreads();
    // Per digit: exclusive prefix over waves, block total, and block-local offset.
    {
        const uint32_t d = threadIdx.x;
        uint32_t tot = 0;
        for (int w = 0; w < kSortWaves; ++w) {
            const uint32_t c = wave_cnt[w][d];
            wave_cnt[w][d] = tot;
            tot += c;
        }
        uint32_t excl;
        block_exclusive_scan(tot, excl, scan_tmp);
        local_off[d] = excl;
        global_base[d] = hist_scanned[size_t(d) * n_blocks + blockIdx.x] - excl;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r) {
        const size_t i = wave0 + size_t(r) * kWave + lane;
        if (i < n) {
            const uint32_t d = (key[r] >> shift) & mask;
            const uint32_t lp = local_off[d] + wave_cnt[wave][d] + rank[r];
            s_keys[lp] = key[r];
            s_vals[lp] = val[r];
        }
    }
    __syncthreads();
    const uint32_t count = uint32_t(min<size_t>(kSortTile, n - tile0));
    for (uint32_t i = threadIdx.x; i < count; i += kSortThreads) {
        const uint32_t k = s_keys[i];
        const uint32_t g = global_base[(k >> shift) & mask] + i;
        keys_out[g] = KOut(k);
        vals_out[g] = s_vals[i];
    }
}

struct SortLayout {
    uint32_t n_blocks;
    size_t off_a_keys, off_a_vals, off_b_keys, off_b_vals, off_hist, off_scan, total;
};

static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

static SortLayout sort_layout(uint64_t n) {
    SortLayout L;
    L.n_blocks = uint32_t((n + kSortTile - 1) / kSortTile);
    size_t o = 0;
    const size_t kb = align256(size_t(n) * 4 + 4);
    L.off_a_keys = o; o += kb;
    L.off_a_vals = o; o += kb;
    L.off_b_keys = o; o += kb;
    L.off_b_vals = o; o += kb;
    const size_t hist_n = size_t(kRadix) * L.n_blocks;
    L.off_hist = o; o += align256((hist_n + 1) * 4);
    L.off_scan = o; o += align256(scan_levels_bytes(hist_n));
    L.total = o;
    return L;
}

size_t sort_workspace_bytes(uint64_t n) { return sort_layout(n).total; }

uint32_t* sort_keys_buffer(void* ws) { return reinterpret_cast<uint32_t*>(static_cast<char*>(ws)); }
uint32_t* sort_vals_buffer(void* ws, uint64_t n) {
    return reinterpret_cast<uint32_t*>(static_cast<char*>(ws) + sort_layout(n).off_a_vals);
}

// Sorts the pairs held in the workspace's A buffers (written by an emit kernel).
// `n` is the element count, or a capacity when the live count is on the device
// (`d_n`); the last pass writes keys as KOut and values to the final outputs.
template <typename KOut>
static int sort_pairs_impl(void* ws, size_t ws_bytes, uint32_t n, const uint32_t* d_n, uint32_t end_bit,
                           KOut* keys_out, uint32_t* vals_out, hipStream_t stream) {
    const SortLayout L = sort_layout(n);
    R2S_REQUIRE(ws_bytes >= L.total, "sort workspace too small: %zu < %zu bytes", ws_bytes, L.total);
    if (n == 0) return R2S_OK;
    char* base = static_cast<char*>(ws);
    uint32_t* ka = reinterpret_cast<uint32_t*>(base + L.off_a_keys);
    uint32_t* va = reinterpret_cast<uint32_t*>(base + L.off_a_vals);
    uint32_t* kb = reinterpret_cast<uint32_t*>(base + L.off_b_keys);
    uint32_t* vb = reinterpret_cast<uint32_t*>(base + L.off_b_vals);
    uint32_t* hist = reinterpret_cast<uint32_t*>(base + L.off_hist);
    void* scan_ws = base + L.off_scan;
    const uint32_t hist_n = uint32_t(kRadix) * L.n_blocks;
    const uint32_t bits = end_bit == 0 ? 1u : end_bit;
    const uint32_t passes = (bits + kRadixBits - 1) / kRadixBits;
    for (uint32_t p = 0; p < passes; ++p) {
        const uint32_t shift = p * kRadixBits;
        const uint32_t nb = std::min<uint32_t>(kRadixBits, bits - shift);
        radix_upsweep_kernel<<<L.n_blocks, kSortThreads, 0, stream>>>(ka, n, d_n, shift, nb, L.n_blocks, hist);
        if (int rc = check_launch("radix_upsweep")) return rc;
        if (int rc = scan_exclusive(hist, hist_n, hist, nullptr, scan_ws, stream)) return rc;
        if (p + 1 == passes) {
            radix_downsweep_kernel<KOut><<<L.n_blocks, kSortThreads, 0, stream>>>(
                ka, va, keys_out, vals_out, n, d_n, shift, nb, L.n_blocks, hist);
        } else {
            radix_downsweep_kernel<uint32_t><<<L.n_blocks, kSortThreads, 0, stream>>>(
                ka, va, kb, vb, n, d_n, shift, nb, L.n_blocks, hist);
            uint32_t* t = ka; ka = kb; kb = t;
            t = va; va = vb; vb = t;
        }
        if (int rc = check_launch("radix_downsweep")) return rc;
    }
    return R2S_OK;
}

int sort_pairs_to_output(void* ws, size_t ws_bytes, uint32_t n, uint32_t end_bit, int64_t* keys_out,
                         int32_t* vals_out, hipStream_t stream) {
    return sort_pairs_impl<int64_t>(ws, ws_bytes, n, nullptr, end_bit, keys_out,
                                    reinterpret_cast<uint32_t*>(vals_out), stream);
}

int sort_pairs_u32(void* ws, size_t ws_bytes, uint32_t capacity, const uint32_t* d_n, uint32_t end_bit,
                   uint32_t* keys_out, uint32_t* vals_out, hipStream_t stream) {
    return sort_pairs_impl<uint32_t>(ws, ws_bytes, capacity, d_n, end_bit, keys_out, vals_out, stream);
}

uint32_t key_end_bit(uint64_t max_key) {
    // ceil(log2(max_key + 1)), as the reference computes its sort width.
    uint32_t b = 0;
    while (b < 64 && (uint64_t(1) << b) < max_key + 1) ++b;
    return b;
}

// offsets[k] = lower_bound(ids, k).
__global__ void __launch_bounds__(256)
offset_encode_kernel(const int64_t* __restrict__ ids, uint32_t n, uint32_t n_keys,
                     int32_t* __restrict__ offsets) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_keys) return;
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (ids[mid] < int64_t(k)) lo = mid + 1; else hi = mid;
    }
    offsets[k] = int32_t(lo);
}

}  // namespace r2s

using namespace r2s;

extern "C" size_t r2s_scan_workspace_bytes(uint64_t n) { return scan_workspace_bytes(n); }

extern "C" size_t r2s_sort_workspace_bytes(uint64_t n_isects, uint32_t end_bit) {
    (void)end_bit;
    return sort_workspace_bytes(n_isects);
}

extern "C" int r2s_isect_tiles_count(int dtype, const void* means2d, const int32_t* radii,
                                     int radii_is_rect, uint32_t C, uint32_t N, uint32_t tile_size,
                                     uint32_t n_tiles_x, uint32_t n_tiles_y,
                                     int32_t* tiles_per_gauss, uint32_t* cum, void* scan_ws,
                                     uint32_t* d_total, r2s_stream_t stream) {
    R2S_REQUIRE(tile_size > 0, "tile_size must be positive");
    const uint64_t total64 = uint64_t(C) * N;
    R2S_REQUIRE(total64 < (1ull << 31), "C*N = %llu exceeds the int32 index range",
                (unsigned long long)total64);
    const uint32_t total = uint32_t(total64);
    return dispatch_dtype(dtype, [&](auto tag) {
        using S = decltype(tag);
        if (total > 0) {
            count_2d_kernel<S><<<ceil_div(total, 256), 256, 0, stream>>>(
                (const S*)means2d, radii, radii_is_rect != 0, total, tile_size, n_tiles_x, n_tiles_y,
                tiles_per_gauss, cum);
            if (int rc = check_launch("isect_count_2d")) return rc;
        }
        return scan_exclusive(cum, total, cum, d_total, scan_ws, stream);
    });
}

extern "C" int r2s_isect_tiles_emit(int dtype, const void* means2d, const int32_t* radii,
                                    int radii_is_rect, uint32_t C, uint32_t N, uint32_t tile_size,
                                    uint32_t n_tiles_x, uint32_t n_tiles_y, const uint32_t* cum,
                                    uint32_t n_isects, int64_t* isect_ids, int32_t* flatten_ids,
                                    void* sort_ws, size_t sort_ws_bytes, r2s_stream_t stream) {
    R2S_REQUIRE(n_isects < (1u << 31), "n_isects exceeds the int32 index range");
    if (n_isects == 0) return R2S_OK;
    R2S_REQUIRE(sort_ws_bytes >= sort_workspace_bytes(n_isects), "sort workspace too small");
    const uint64_t max_key = uint64_t(C) * n_tiles_x * n_tiles_y;
    R2S_REQUIRE(max_key < (1ull << 32), "C*n_tiles exceeds the 32-bit key range");
    const uint32_t total = C * N;
    return dispatch_dtype(dtype, [&](auto tag) {
        using S = decltype(tag);
        emit_2d_kernel<S><<<ceil_div(total, 256), 256, 0, stream>>>(
            (const S*)means2d, radii, radii_is_rect != 0, total, N, tile_size, n_tiles_x, n_tiles_y,
            cum, sort_keys_buffer(sort_ws), sort_vals_buffer(sort_ws, n_isects));
        if (int rc = check_launch("isect_emit_2d")) return rc;
        return sort_pairs_to_output(sort_ws, sort_ws_bytes, n_isects, key_end_bit(max_key), isect_ids,
                                    flatten_ids, stream);
    });
}

extern "C" int r2s_isect_tiles_3d_count(int dtype, const void* means3d, const int32_t* radii,
                                        uint32_t N, uint32_t tile_size, uint32_t n_tiles_x,
                                        uint32_t n_tiles_y, uint32_t n_tiles_z,
                                        int32_t* tiles_per_gauss, uint32_t* cum, void* scan_ws,
                                        uint32_t* d_total, r2s_stream_t stream) {
    R2S_REQUIRE(tile_size > 0, "tile_size must be positive");
    R2S_REQUIRE(N < (1u << 31), "N exceeds the int32 index range");
    return dispatch_dtype(dtype, [&](auto tag) {
        using S = decltype(tag);
        if (N > 0) {
            count_3d_kernel<S><<<ceil_div(N, 256), 256, 0, stream>>>(
                (const S*)means3d, radii, N, tile_size, n_tiles_x, n_tiles_y, n_tiles_z,
                tiles_per_gauss, cum);
            if (int rc = check_launch("isect_count_3d")) return rc;
        }
        return scan_exclusive(cum, N, cum, d_total, scan_ws, stream);
    });
}

extern "C" int r2s_isect_tiles_3d_emit(int dtype, const void* means3d, const int32_t* radii,
                                       uint32_t N, uint32_t tile_size, uint32_t n_tiles_x,
                                       uint32_t n_tiles_y, uint32_t n_tiles_z, const uint32_t* cum,
                                       uint32_t n_isects, int64_t* isect_ids, int32_t* flatten_ids,
                                       void* sort_ws, size_t sort_ws_bytes, r2s_stream_t stream) {
    R2S_REQUIRE(n_isects < (1u << 31), "n_isects exceeds the int32 index range");
    if (n_isects == 0) return R2S_OK;
    R2S_REQUIRE(sort_ws_bytes >= sort_workspace_bytes(n_isects), "sort workspace too small");
    const uint64_t max_key = uint64_t(n_tiles_x) * n_tiles_y * n_tiles_z;
    R2S_REQUIRE(max_key < (1ull << 32), "n_tiles exceeds the 32-bit key range");
    return dispatch_dtype(dtype, [&](auto tag) {
        using S = decltype(tag);
        emit_3d_kernel<S><<<ceil_div(N, 256), 256, 0, stream>>>(
            (const S*)means3d, radii, N, tile_size, n_tiles_x, n_tiles_y, n_tiles_z, cum,
            sort_keys_buffer(sort_ws), sort_vals_buffer(sort_ws, n_isects));
        if (int rc = check_launch("isect_emit_3d")) return rc;
        return sort_pairs_to_output(sort_ws, sort_ws_bytes, n_isects, key_end_bit(max_key), isect_ids,
                                    flatten_ids, stream);
    });
}

extern "C" int r2s_isect_offset_encode(const int64_t* isect_ids, uint32_t n_isects, uint32_t n_keys,
                                       int32_t* offsets, r2s_stream_t stream) {
    if (n_keys == 0) return R2S_OK;
    offset_encode_kernel<<<ceil_div(n_keys, 256), 256, 0, stream>>>(isect_ids, n_isects, n_keys, offsets);
    return check_launch("isect_offset_encode");
}

// Internal interface of the binning primitives (scan, stable radix sort).
#pragma once

#include "common.hpp"

namespace r2s {

// Tile range of one axis (IsectTilesKernel.cu:30-40,91-103): floor/ceil of
// (mean -/+ radius) / tile, clamped to [0, n_tiles] in floating point before
// the integer conversion. Returns false when the range is empty.
template <typename S>
__device__ __forceinline__ uint32_t clamp_bound(S v, uint32_t upper) {
    if (!isfinite(v) || v <= S(0)) return 0u;
    if (v >= S(upper)) return upper;
    return static_cast<uint32_t>(v);
}

template <typename S>
__device__ __forceinline__ bool tile_span(S mean, int32_t radius, uint32_t tile, uint32_t n_tiles,
                                          uint32_t& lo, uint32_t& hi) {
    const S t = S(tile);
    const S c = mean / t, r = S(radius) / t;
    lo = clamp_bound(floor(c - r), n_tiles);
    hi = clamp_bound(ceil(c + r), n_tiles);
    return hi > lo;
}

// Tile box of one 3D primitive (IsectTiles3DKernel.cu:36-45): per axis the
// tile_span of mean +/- radius; false when a radius is not positive or a span
// is empty. Emission order within the box is z, then y, then x.
template <typename S>
__device__ __forceinline__ bool box_3d(const S* __restrict__ means3d, const int32_t* __restrict__ radii,
                                       size_t n, uint32_t tile, uint32_t ntx, uint32_t nty,
                                       uint32_t ntz, uint32_t lo[3], uint32_t hi[3]) {
    const int32_t r[3] = {radii[3 * n], radii[3 * n + 1], radii[3 * n + 2]};
    if (r[0] <= 0 || r[1] <= 0 || r[2] <= 0) return false;
    const uint32_t nt[3] = {ntx, nty, ntz};
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) ok = tile_span(means3d[3 * n + k], r[k], tile, nt[k], lo[k], hi[k]) && ok;
    return ok;
}

size_t scan_workspace_bytes(uint64_t n);
// Exclusive scan of n uint32 values: out[0..n], out[n] = total (also written to
// *d_total when non-null). `out` may alias `in`; `out` must hold n + 1 values.
int scan_exclusive(const uint32_t* in, uint32_t n, uint32_t* out, uint32_t* d_total, void* ws,
                   hipStream_t stream);

size_t sort_workspace_bytes(uint64_t n);
uint32_t* sort_keys_buffer(void* ws);
uint32_t* sort_vals_buffer(void* ws, uint64_t n);
uint32_t key_end_bit(uint64_t max_key);
int sort_pairs_to_output(void* ws, size_t ws_bytes, uint32_t n, uint32_t end_bit, int64_t* keys_out,
                         int32_t* vals_out, hipStream_t stream);
// Capacity-bounded variant: sorts min(*d_n, capacity) pairs, grids sized for
// `capacity`, so the launch sequence needs no host-side count.
int sort_pairs_u32(void* ws, size_t ws_bytes, uint32_t capacity, const uint32_t* d_n, uint32_t end_bit,
                   uint32_t* keys_out, uint32_t* vals_out, hipStream_t stream);

}  // namespace r2s

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

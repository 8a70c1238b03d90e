// Internal interface of the binning primitives (scan, stable radix sort).
#pragma once

#include "common.hpp"

namespace r2s {

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

}  // namespace r2s

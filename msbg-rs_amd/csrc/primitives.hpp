// Device-wide primitives shared by the particle kernels: exclusive scan,
// stable LSD radix sort of (u32 key, u32 value) pairs, and a grow-only
// scratch workspace.  All run on the library stream.
#pragma once

#include "internal.hpp"

namespace msbg {

// Grow-only device workspace, one per slot id (not thread safe: one host
// thread drives the library per device).
void *workspace(int id, size_t bytes, int *rc);

// out[i] = sum_{j<i} in[i]; returns the total in *total_host if non-null
// (synchronizes only then).  in and out may alias.
int scan_exclusive_u32(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *total_host);
// The same scan with the caller's tile-sum buffer of scan_sums_words(n) words
// (no workspace allocation).
int scan_exclusive_u32_with(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *sums, uint32_t *total_host);
int64_t scan_sums_words(int64_t n);

// Stable sort of (keys, vals) by the low `key_bits` bits of keys, in place.
int radix_sort_pairs(uint32_t *keys, uint32_t *vals, int64_t n, int key_bits);

// msbg_morton_sort without the final synchronization (stream-ordered).
int morton_sort_async(const float *d_pos, int64_t n, uint32_t *d_keys, uint32_t *d_perm);
// morton_sort_async, also writing each particle's records (x, y, z, m), (u, v, w, 0)
// to d_rec[2 i], d_rec[2 i + 1] in caller order (d_mass null: m = 1).
int morton_sort_pack_async(const float *d_pos, const float *d_vel, const float *d_mass, int64_t n, uint32_t *d_keys,
                           uint32_t *d_perm, float4 *d_rec);

__device__ __forceinline__ uint32_t spread10(uint32_t v) {  // multires/sort.rs:8-16, 10 bits
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// Rust `f32 as usize`: truncation, saturating at 0 for negatives and NaN.
__device__ __forceinline__ uint32_t trunc_index(float p) {
    return p > 0.0f ? (p < 4294967040.0f ? (uint32_t)p : 0xffffffffu) : 0u;
}

}  // namespace msbg

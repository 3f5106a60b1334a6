// Particle-side kernels: Morton keys, the Morton sort, block bucketing
// (bucket_indices) and active-block compaction (active_blocks).
#include <algorithm>
#include <mutex>

#include "primitives.hpp"

namespace msbg {

// morton3 (multires/sort.rs:20) of the particle's voxel (trunc position,
// p2g.rs:153), low 30 bits: exact for voxel coordinates below 1024.
__global__ void morton_keys_kernel(const float *__restrict__ pos, int64_t n, uint32_t *__restrict__ keys,
                                   uint32_t *__restrict__ idx) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t x = trunc_index(pos[3 * i]), y = trunc_index(pos[3 * i + 1]), z = trunc_index(pos[3 * i + 2]);
        keys[i] = spread10(x) | (spread10(y) << 1) | (spread10(z) << 2);
        if (idx) idx[i] = (uint32_t)i;
    }
}

// morton_keys_kernel's keys and indices, and each particle's two 16-byte
// records (x, y, z, m) and (u, v, w, 0) in caller order, so that the sorted
// payload is gathered as one 32-byte record per particle instead of three
// scattered reads.
__global__ void morton_keys_pack_kernel(const float *__restrict__ pos, const float *__restrict__ vel,
                                        const float *__restrict__ mass, int64_t n, uint32_t *__restrict__ keys,
                                        uint32_t *__restrict__ idx, float4 *__restrict__ rec) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
        keys[i] = spread10(trunc_index(px)) | (spread10(trunc_index(py)) << 1) | (spread10(trunc_index(pz)) << 2);
        idx[i] = (uint32_t)i;
        rec[2 * i] = make_float4(px, py, pz, mass ? mass[i] : 1.0f);
        rec[2 * i + 1] = make_float4(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], 0.0f);
    }
}

// Sorted particle arrays from the packed records: position, velocity and
// mass of particle perm[i], one 32-byte record read per particle.
__global__ void gather_particles_kernel(const float4 *__restrict__ rec, const uint32_t *__restrict__ perm, int64_t n,
                                        float *__restrict__ pos, float *__restrict__ vel, float *__restrict__ mass) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = perm[i];
        const float4 a = rec[2 * j], b = rec[2 * j + 1];
        pos[3 * i] = a.x;
        pos[3 * i + 1] = a.y;
        pos[3 * i + 2] = a.z;
        vel[3 * i] = b.x;
        vel[3 * i + 1] = b.y;
        vel[3 * i + 2] = b.z;
        mass[i] = a.w;
    }
}

// Row-major block id of a particle (p2g.rs:149-157).
__global__ void particle_bids_kernel(const float *__restrict__ pos, int64_t n, int32_t nx, int32_t ny,
                                     uint32_t *__restrict__ bids, uint32_t *__restrict__ idx) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t x = trunc_index(pos[3 * i]), y = trunc_index(pos[3 * i + 1]), z = trunc_index(pos[3 * i + 2]);
        bids[i] = (x >> 4) + (y >> 4) * (uint32_t)nx + (z >> 4) * (uint32_t)nx * (uint32_t)ny;
        if (idx) idx[i] = (uint32_t)i;
    }
}

__global__ void iota_kernel(uint32_t *v, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        v[i] = (uint32_t)i;
}

__global__ void run_heads_kernel(const uint32_t *__restrict__ keys, int64_t n, uint32_t *__restrict__ flags) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

__global__ void run_write_kernel(const uint32_t *__restrict__ keys, int64_t n, const uint32_t *__restrict__ heads,
                                 const uint32_t *__restrict__ rank, int32_t *__restrict__ blocks,
                                 int32_t *__restrict__ starts) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (heads[i]) {
            blocks[rank[i]] = (int32_t)keys[i];
            starts[rank[i]] = (int32_t)i;
        }
}

__global__ void set_i32_kernel(int32_t *p, int32_t v) { *p = v; }

__global__ void iota_fill_kernel(uint32_t *p, int64_t n, uint32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// footprint_axis (particles/mod.rs:207-212): trunc(p/bsx -+ r_scan/bsx) clipped.
__device__ __forceinline__ void footprint(float p, float rb, int32_t nb, int32_t &lo, int32_t &hi) {
    const float b = p / 16.0f;
    lo = max((int32_t)truncf(b - rb), 0);
    hi = min((int32_t)truncf(b + rb), nb - 1);
}

// active_blocks (particles/active.rs:15-47): every particle marks its footprint
// blocks (a benign race: all writers store 1).
__global__ void mark_active_kernel(const float *__restrict__ pos, int64_t n, float rb, int32_t nx, int32_t ny,
                                   int32_t nz, uint32_t *__restrict__ flags) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int32_t x1, x2, y1, y2, z1, z2;
        footprint(pos[3 * i], rb, nx, x1, x2);
        footprint(pos[3 * i + 1], rb, ny, y1, y2);
        footprint(pos[3 * i + 2], rb, nz, z1, z2);
        for (int32_t bz = z1; bz <= z2; ++bz)
            for (int32_t by = y1; by <= y2; ++by)
                for (int32_t bx = x1; bx <= x2; ++bx) flags[bx + (int64_t)by * nx + (int64_t)bz * nx * ny] = 1u;
    }
}

__global__ void compact_flags_kernel(const uint32_t *__restrict__ flags, const uint32_t *__restrict__ rank,
                                     int64_t n, int32_t *__restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (flags[i]) out[rank[i]] = (int32_t)i;
}

template <typename T>
__global__ void gather_kernel_rows(const T *__restrict__ src, const uint32_t *__restrict__ perm, int64_t n, int w,
                                   T *__restrict__ dst) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n * w; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = i / w, c = i - r * w;
        dst[i] = src[(int64_t)perm[r] * w + c];
    }
}

unsigned grid_for(int64_t n, int threads = 256) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + threads - 1) / threads, 65536));
}

// Occupied blocks and CSR starts of sorted keys; *n_occ is read back.
int run_lengths(const uint32_t *keys, int64_t n, int32_t *blocks, int32_t *starts, int64_t *n_occ) {
    int rc;
    uint32_t *heads = (uint32_t *)workspace(3, sizeof(uint32_t) * 2 * std::max<int64_t>(n, 1), &rc);
    if (rc) return rc;
    uint32_t *rank = heads + n;
    hipLaunchKernelGGL(run_heads_kernel, dim3(grid_for(n)), dim3(256), 0, stream(), keys, n, heads);
    MSBG_HIP(hipGetLastError());
    uint32_t total = 0;
    rc = scan_exclusive_u32(heads, rank, n, &total);
    if (rc) return rc;
    hipLaunchKernelGGL(run_write_kernel, dim3(grid_for(n)), dim3(256), 0, stream(), keys, n, heads, rank, blocks,
                       starts);
    hipLaunchKernelGGL(set_i32_kernel, dim3(1), dim3(1), 0, stream(), starts + total, (int32_t)n);
    MSBG_HIP(hipGetLastError());
    MSBG_HIP(hipStreamSynchronize(stream()));
    *n_occ = total;
    return MSBG_OK;
}

// Marks, scans and compacts the footprint blocks of n particles into d_out
// (ascending); flags / rank hold nb words each, sums scan_sums_words(nb).
// The flags are cleared by a fill kernel, not hipMemsetAsync, so the whole
// sequence is plain kernel launches plus the count's read-back.
int active_blocks_into(const float *d_pos, int64_t n, float r_scan, int32_t nx, int32_t ny, int32_t nz,
                       uint32_t *flags, uint32_t *rank, uint32_t *sums, int32_t *d_out, int64_t *n_out) {
    const int64_t nb = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(iota_fill_kernel, dim3(grid_for(nb)), dim3(256), 0, stream(), flags, nb, 0u);
    if (n > 0)
        hipLaunchKernelGGL(mark_active_kernel, dim3(grid_for(n)), dim3(256), 0, stream(), d_pos, n, r_scan / 16.0f, nx,
                           ny, nz, flags);
    MSBG_HIP(hipGetLastError());
    uint32_t total = 0;
    int rc = scan_exclusive_u32_with(flags, rank, nb, sums, &total);
    if (rc) return rc;
    hipLaunchKernelGGL(compact_flags_kernel, dim3(grid_for(nb)), dim3(256), 0, stream(), flags, rank, nb, d_out);
    MSBG_HIP(hipGetLastError());
    MSBG_HIP(hipStreamSynchronize(stream()));
    *n_out = total;
    return MSBG_OK;
}

// ---------------------------------------------------------------------------
// Fixed scratch of msbg_active_blocks_noalloc: the flags and ranks of up to
// kNoAllocMaxBlocks blocks and the scan's tile sums, one copy per GPU for the
// life of the process (2.1 MB).  Host code reaches it only through the
// address hipGetSymbolAddress returns for the current device.
// ---------------------------------------------------------------------------
constexpr int64_t kNoAllocMaxBlocks = MSBG_NOALLOC_MAX_BLOCKS;
constexpr int64_t kNoAllocSumWords = 1024;  // >= scan_sums_words(kNoAllocMaxBlocks) = 129
constexpr int64_t kNoAllocWords = 2 * kNoAllocMaxBlocks + kNoAllocSumWords;
static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static uint32_t *g_scratch[kMaxDevices] = {nullptr};

// The current device's copy of `scratch`, resolved on its first use there.
int noalloc_scratch(uint32_t **out) {
    int dev = 0;
    MSBG_HIP(hipGetDevice(&dev));
    MSBG_CHECK(dev >= 0 && dev < kMaxDevices, MSBG_EINVAL, "no-alloc scratch: device index beyond 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        MSBG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch[dev] = (uint32_t *)ptr;
    }
    *out = g_scratch[dev];
    return MSBG_OK;
}

int morton_sort_async(const float *d_pos, int64_t n, uint32_t *d_keys, uint32_t *d_perm) {
    if (n == 0) return MSBG_OK;
    hipLaunchKernelGGL(morton_keys_kernel, dim3(grid_for(n)), dim3(256), 0, stream(), d_pos, n, d_keys, d_perm);
    MSBG_HIP(hipGetLastError());
    return radix_sort_pairs(d_keys, d_perm, n, 30);
}

int morton_sort_pack_async(const float *d_pos, const float *d_vel, const float *d_mass, int64_t n, uint32_t *d_keys,
                           uint32_t *d_perm, float4 *d_rec) {
    if (n == 0) return MSBG_OK;
    hipLaunchKernelGGL(morton_keys_pack_kernel, dim3(grid_for(n)), dim3(256), 0, stream(), d_pos, d_vel, d_mass, n,
                       d_keys, d_perm, d_rec);
    MSBG_HIP(hipGetLastError());
    return radix_sort_pairs(d_keys, d_perm, n, 30);
}

int bits_for(int64_t v) {
    int b = 1;
    while (b < 32 && ((int64_t)1 << b) <= v) ++b;
    return b;
}

}  // namespace msbg

using namespace msbg;

extern "C" {

int msbg_morton_keys(const float *d_pos, int64_t n, uint32_t *d_keys) {
    MSBG_CHECK(n >= 0 && (n == 0 || (d_pos && d_keys)), MSBG_EINVAL, "morton keys: null argument");
    if (n == 0) return MSBG_OK;
    hipLaunchKernelGGL(morton_keys_kernel, dim3(grid_for(n)), dim3(256), 0, stream(), d_pos, n, d_keys, nullptr);
    MSBG_HIP(hipGetLastError());
    MSBG_HIP(hipStreamSynchronize(stream()));
    return MSBG_OK;
}

int msbg_radix_sort_pairs(uint32_t *d_keys, uint32_t *d_vals, int64_t n, int key_bits) {
    MSBG_CHECK(n == 0 || (d_keys && d_vals), MSBG_EINVAL, "radix sort: null argument");
    int rc = radix_sort_pairs(d_keys, d_vals, n, key_bits);
    if (rc) return rc;
    MSBG_HIP(hipStreamSynchronize(stream()));
    return MSBG_OK;
}

int msbg_morton_sort(const float *d_pos, int64_t n, uint32_t *d_keys, uint32_t *d_perm) {
    MSBG_CHECK(n >= 0 && n < ((int64_t)1 << 31), MSBG_EINVAL, "morton sort: n must be below 2^31");
    MSBG_CHECK(n == 0 || (d_pos && d_keys && d_perm), MSBG_EINVAL, "morton sort: null argument");
    if (n == 0) return MSBG_OK;
    int rc = morton_sort_async(d_pos, n, d_keys, d_perm);
    if (rc) return rc;
    MSBG_HIP(hipStreamSynchronize(stream()));
    return MSBG_OK;
}

int msbg_particle_bids(int64_t sx, int64_t sy, int64_t sz, const float *d_pos, int64_t n, int32_t *d_bids) {
    MSBG_CHECK(sx > 0 && sy > 0 && sz > 0 && n >= 0 && (n == 0 || (d_pos && d_bids)), MSBG_EINVAL,
               "particle bids: bad argument");
    if (n == 0) return MSBG_OK;
    const int32_t nx = (int32_t)((sx + 15) >> 4), ny = (int32_t)((sy + 15) >> 4);
    hipLaunchKernelGGL(particle_bids_kernel, dim3(grid_for(n)), dim3(256), 0, stream(), d_pos, n, nx, ny,
                       (uint32_t *)d_bids, nullptr);
    MSBG_HIP(hipGetLastError());
    MSBG_HIP(hipStreamSynchronize(stream()));
    return MSBG_OK;
}

int msbg_bucket_indices(const int32_t *d_bids, int64_t n, int64_t n_blocks, int32_t *d_perm,
                        int32_t *d_particle_blocks, int32_t *d_starts, int64_t *n_occupied) {
    MSBG_CHECK(n_occupied && d_starts, MSBG_EINVAL, "bucket_indices: null output");
    MSBG_CHECK(n >= 0 && n < ((int64_t)1 << 31) && n_blocks > 0 && n_blocks < ((int64_t)1 << 31), MSBG_EINVAL,
               "bucket_indices: n and n_blocks must be below 2^31");
    MSBG_CHECK(n == 0 || (d_bids && d_perm && d_particle_blocks), MSBG_EINVAL, "bucket_indices: null argument");
    if (n == 0) {  // particles/sort.rs:176-181
        hipLaunchKernelGGL(set_i32_kernel, dim3(1), dim3(1), 0, stream(), d_starts, 0);
        MSBG_HIP(hipGetLastError());
        MSBG_HIP(hipStreamSynchronize(stream()));
        *n_occupied = 0;
        return MSBG_OK;
    }
    int rc;
    uint32_t *keys = (uint32_t *)workspace(4, sizeof(uint32_t) * n, &rc);
    if (rc) return rc;
    MSBG_HIP(hipMemcpyAsync(keys, d_bids, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, stream()));
    hipLaunchKernelGGL(iota_kernel, dim3(grid_for(n)), dim3(256), 0, stream(), (uint32_t *)d_perm, n);
    MSBG_HIP(hipGetLastError());
    rc = radix_sort_pairs(keys, (uint32_t *)d_perm, n, bits_for(n_blocks - 1));
    if (rc) return rc;
    return run_lengths(keys, n, d_particle_blocks, d_starts, n_occupied);
}

int msbg_active_blocks(int64_t sx, int64_t sy, int64_t sz, const float *d_pos, int64_t n, float r_scan,
                       int32_t *d_out, int64_t *n_out) {
    MSBG_CHECK(n_out && d_out, MSBG_EINVAL, "active_blocks: null output");
    MSBG_CHECK(sx > 0 && sy > 0 && sz > 0 && n >= 0 && (n == 0 || d_pos), MSBG_EINVAL, "active_blocks: bad argument");
    const int32_t nx = (int32_t)((sx + 15) >> 4), ny = (int32_t)((sy + 15) >> 4), nz = (int32_t)((sz + 15) >> 4);
    const int64_t nb = (int64_t)nx * ny * nz;
    MSBG_CHECK(nb < ((int64_t)1 << 31), MSBG_EINVAL, "active_blocks: block grid too large for a dense flag array");
    int rc;
    uint32_t *flags = (uint32_t *)workspace(5, sizeof(uint32_t) * 2 * nb, &rc);
    if (rc) return rc;
    uint32_t *sums = (uint32_t *)workspace(0, sizeof(uint32_t) * scan_sums_words(nb), &rc);
    if (rc) return rc;
    return active_blocks_into(d_pos, n, r_scan, nx, ny, nz, flags, flags + nb, sums, d_out, n_out);
}

int msbg_active_blocks_noalloc(int64_t sx, int64_t sy, int64_t sz, const float *d_pos, int64_t n, float r_scan,
                               int32_t *d_out, int64_t *n_out) {
    MSBG_CHECK(n_out && d_out, MSBG_EINVAL, "active_blocks_noalloc: null output");
    MSBG_CHECK(sx > 0 && sy > 0 && sz > 0 && n >= 0 && (n == 0 || d_pos), MSBG_EINVAL,
               "active_blocks_noalloc: bad argument");
    const int64_t nx = (sx + 15) >> 4, ny = (sy + 15) >> 4, nz = (sz + 15) >> 4;
    MSBG_CHECK(nx * ny * nz <= kNoAllocMaxBlocks, MSBG_EINVAL,
               "active_blocks_noalloc: the block grid exceeds MSBG_NOALLOC_MAX_BLOCKS (262,144 blocks, a 1024^3 "
               "domain of 16^3 blocks); use msbg_active_blocks");
    static_assert(kNoAllocSumWords >= (kNoAllocMaxBlocks + 2047) / 2048 + 1, "scan tile sums do not fit");
    uint32_t *base = nullptr;
    int rc = noalloc_scratch(&base);
    if (rc) return rc;
    const int64_t nb = nx * ny * nz;
    MSBG_CHECK(scan_sums_words(nb) <= kNoAllocSumWords, MSBG_EINVAL, "active_blocks_noalloc: scan sums overflow");
    return active_blocks_into(d_pos, n, r_scan, (int32_t)nx, (int32_t)ny, (int32_t)nz, base, base + kNoAllocMaxBlocks,
                              base + 2 * kNoAllocMaxBlocks, d_out, n_out);
}

int msbg_morton_sort_particles(const float *d_pos, const float *d_vel, const float *d_mass, int64_t n,
                               uint32_t *d_keys, uint32_t *d_perm, float *d_pos_out, float *d_vel_out,
                               float *d_mass_out) {
    MSBG_CHECK(n >= 0 && (n == 0 || (d_pos && d_vel && d_keys && d_perm && d_pos_out && d_vel_out && d_mass_out)),
               MSBG_EINVAL, "morton sort particles: null argument");
    MSBG_CHECK(n < ((int64_t)1 << 32), MSBG_EINVAL, "morton sort particles: more than 2^32 particles");
    if (n == 0) return MSBG_OK;
    int rc = 0;
    float4 *rec = (float4 *)workspace(22, sizeof(float4) * 2 * (size_t)n, &rc);
    if (rc) return rc;
    rc = morton_sort_pack_async(d_pos, d_vel, d_mass, n, d_keys, d_perm, rec);
    if (rc) return rc;
    hipLaunchKernelGGL(gather_particles_kernel, dim3(grid_for(n)), dim3(256), 0, stream(), rec, d_perm, n, d_pos_out,
                       d_vel_out, d_mass_out);
    MSBG_HIP(hipGetLastError());
    MSBG_HIP(hipStreamSynchronize(stream()));
    return MSBG_OK;
}

int msbg_gather_rows(const void *d_src, const uint32_t *d_perm, int64_t n, int64_t row_bytes, void *d_dst) {
    MSBG_CHECK(n >= 0 && (n == 0 || (d_src && d_perm && d_dst)), MSBG_EINVAL, "gather rows: null argument");
    MSBG_CHECK(row_bytes > 0 && row_bytes % 4 == 0, MSBG_EINVAL, "gather rows: row_bytes must be a multiple of 4");
    if (n == 0) return MSBG_OK;
    const int w = (int)(row_bytes / 4);
    hipLaunchKernelGGL(gather_kernel_rows<uint32_t>, dim3(grid_for(n * w)), dim3(256), 0, stream(),
                       (const uint32_t *)d_src, d_perm, n, w, (uint32_t *)d_dst);
    MSBG_HIP(hipGetLastError());
    MSBG_HIP(hipStreamSynchronize(stream()));
    return MSBG_OK;
}

}  // extern "C"

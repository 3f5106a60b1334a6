// Density min-splat: particles::splat (particles/splat.rs:92-319) for
// D = Density (u16, channel.rs:303-375) into a u16 slab over a grid layout.
//
// A block's particle bucket is cut into work items of at most
// kItemParticles particles, and each item's (16 + 2m)^3 window
// (m = ceil(r_scan)) into kZSlabs z-slabs, one work-group each: the
// work-group stages its slab in LDS as u32 cells initialised to 65535, walks
// the item's particles in chunks (positions and footprint ranges staged in
// LDS, keeping the particles whose footprint reaches the slab), lowers every
// (particle, footprint row) pair's cells with an LDS atomic min of the
// quantised distance, then commits its slab to the blocks it covers with
// dst = min(dst, staging).  Buckets run in 8 colour passes (block-coordinate
// parity), so same-colour windows never overlap (16 + 2m <= 32); the slabs of
// one item are disjoint, and only the items of one split block (a dense
// point-cloud instance) commit to the same cells, with an atomic min.  min is
// associative and commutative, so the result does not depend on any order.
#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.hpp"
#include "primitives.hpp"

namespace msbg {
namespace {

struct SplatParams {
    int sx, sy, sz, nx, ny;
    int m, W;
    float r, dsq_max, inv;
};

constexpr int kMaxMargin = 6;   // window edge W <= 28
constexpr int kZSlabs = 8;      // work-groups per window
constexpr int kSlabPlanes = 4;  // ceil(28 / 8)
constexpr int kChunk = 256;     // particles staged per pass (one per thread)

// A work item of one colour pass: particles p0 .. p1 - 1 (bucketed order) of
// occupied block blocks[k]; `shared` when the block's particles are split
// over several items, whose windows then commit with an atomic min.
struct DItem {
    int32_t k, p0, p1, shared;
};
constexpr int kItemParticles = 1024;  // particles per item: a dense block (one point-cloud instance) is split

// dst = min(dst, v) on a 16-bit cell, atomically (compare-and-swap on its 32-bit word).
__device__ __forceinline__ void atomic_min_u16(uint16_t *p, uint32_t v) {
    uint32_t *w = reinterpret_cast<uint32_t *>(reinterpret_cast<uintptr_t>(p) & ~(uintptr_t)3);
    const int sh = (reinterpret_cast<uintptr_t>(p) & 2) ? 16 : 0;
    uint32_t old = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (((old >> sh) & 0xffffu) > v) {
        const uint32_t want = (old & ~(0xffffu << sh)) | (v << sh);
        const uint32_t seen = atomicCAS(w, old, want);
        if (seen == old) break;
        old = seen;
    }
}

__global__ __launch_bounds__(256) void density_splat_kernel(SplatParams P, const int32_t *__restrict__ b2s,
                                                            uint16_t *slab, const float *__restrict__ pos,
                                                            const int32_t *__restrict__ blocks,
                                                            const DItem *__restrict__ items) {
    __shared__ uint32_t st[28 * 28 * kSlabPlanes];
    __shared__ int rslot[27];
    __shared__ float s_p[kChunk][3];      // particle position - 0.5
    __shared__ int16_t s_r[kChunk][6];    // footprint: ix1, ix2, iy1, iy2, iz1, iz2 (window coordinates, slab-clipped in z)
    __shared__ int s_n;
    const DItem item = items[blockIdx.x / kZSlabs];
    const int zs = blockIdx.x % kZSlabs;
    const int bid = blocks[item.k];
    const int bx = bid % P.nx, by = (bid / P.nx) % P.ny, bz = bid / (P.nx * P.ny);
    const int W = P.W, m = P.m;
    const int z0 = zs * W / kZSlabs, z1 = (zs + 1) * W / kZSlabs;  // window planes [z0, z1) of this work-group
    const int cells = W * W * (z1 - z0);
    if (cells <= 0) return;
    for (int i = threadIdx.x; i < cells; i += blockDim.x) st[i] = 0xffffu;
    const int b0x = bx * 16, b0y = by * 16, b0z = bz * 16, hi = 16 + m - 1;
    if (threadIdx.x < 27) {
        // region (rx, ry, rz): skipped when any voxel lies outside the domain
        // or the target block is not a value block (splat.rs:275-292)
        const int t = threadIdx.x, rx = t % 3 - 1, ry = (t / 3) % 3 - 1, rz = t / 9 - 1;
        const int gx0 = b0x + (rx < 0 ? -m : (rx == 0 ? 0 : 16)), gy0 = b0y + (ry < 0 ? -m : (ry == 0 ? 0 : 16)),
                  gz0 = b0z + (rz < 0 ? -m : (rz == 0 ? 0 : 16));
        const int wx = rx == 0 ? 16 : m, wy = ry == 0 ? 16 : m, wz = rz == 0 ? 16 : m;
        int slot = -1;
        if (gx0 >= 0 && gy0 >= 0 && gz0 >= 0 && gx0 + wx <= P.sx && gy0 + wy <= P.sy && gz0 + wz <= P.sz) {
            const int s = b2s[(gx0 >> 4) + (gy0 >> 4) * P.nx + (gz0 >> 4) * P.nx * P.ny];
            if (s >= 2) slot = s;
        }
        rslot[t] = slot;
    }
    // stage (splat.rs:179-225): footprint ceil(p - r - 0.5) .. floor(p + r - 0.5),
    // clipped to the window; d2 = dx*dx + (dz*dz + dy*dy)
    const int p0 = item.p0, p1 = item.p1;
    const int side = 2 * m + 1;
    for (int c0 = p0; c0 < p1; c0 += kChunk) {
        if (threadIdx.x == 0) s_n = 0;
        __syncthreads();
        const int j = c0 + (int)threadIdx.x;
        if (j < p1) {
            const float px = pos[3 * j], py = pos[3 * j + 1], pz = pos[3 * j + 2];
            const int iz1 = max(min(max((int)ceilf((pz - P.r) - 0.5f), b0z - m), b0z + hi) - (b0z - m), z0);
            const int iz2 = min(min(max((int)floorf((pz + P.r) - 0.5f), b0z - m), b0z + hi) - (b0z - m), z1 - 1);
            if (iz1 <= iz2) {
                const int q = atomicAdd(&s_n, 1);
                s_p[q][0] = px - 0.5f;
                s_p[q][1] = py - 0.5f;
                s_p[q][2] = pz - 0.5f;
                s_r[q][0] = (int16_t)(min(max((int)ceilf((px - P.r) - 0.5f), b0x - m), b0x + hi) - (b0x - m));
                s_r[q][1] = (int16_t)(min(max((int)floorf((px + P.r) - 0.5f), b0x - m), b0x + hi) - (b0x - m));
                s_r[q][2] = (int16_t)(min(max((int)ceilf((py - P.r) - 0.5f), b0y - m), b0y + hi) - (b0y - m));
                s_r[q][3] = (int16_t)(min(max((int)floorf((py + P.r) - 0.5f), b0y - m), b0y + hi) - (b0y - m));
                s_r[q][4] = (int16_t)iz1;
                s_r[q][5] = (int16_t)iz2;
            }
        }
        __syncthreads();
        const int n = s_n, rows = kSlabPlanes * side;
        for (int i = threadIdx.x; i < n * rows; i += blockDim.x) {
            const int q = i / rows, rr = i - q * rows;
            const int cz = s_r[q][4] + rr / side, cy = s_r[q][2] + rr % side;
            if (cz > s_r[q][5] || cy > s_r[q][3]) continue;
            // window cell (cx, cy, cz) is grid voxel (b0 - m) + c
            const float dz = (float)(b0z - m + cz) - s_p[q][2], dy = (float)(b0y - m + cy) - s_p[q][1];
            const float dyz = dz * dz + dy * dy;
            uint32_t *row = st + ((cz - z0) * W + cy) * W;
            const int cx2 = s_r[q][1];
            for (int cx = s_r[q][0]; cx <= cx2; ++cx) {
                const float dx = (float)(b0x - m + cx) - s_p[q][0];
                const float d2 = dx * dx + dyz;
                if (d2 <= P.dsq_max) atomicMin(row + cx, (uint32_t)(int)roundf((d2 * P.inv) * 65535.0f));
            }
        }
        __syncthreads();
    }
    __syncthreads();  // the staging and the region table are complete even for an empty bucket
    // commit (splat.rs:227-313): each touched window cell, dst = min(dst, staged)
    for (int c = threadIdx.x; c < cells; c += blockDim.x) {
        const uint32_t v = st[c];
        if (v == 0xffffu) continue;
        const int cx = c % W, cy = (c / W) % W, cz = z0 + c / (W * W);
        const int rx = cx < m ? 0 : (cx < m + 16 ? 1 : 2), ry = cy < m ? 0 : (cy < m + 16 ? 1 : 2),
                  rz = cz < m ? 0 : (cz < m + 16 ? 1 : 2);
        const int slot = rslot[rx + 3 * ry + 9 * rz];
        if (slot < 0) continue;
        const int gx = b0x + cx - m, gy = b0y + cy - m, gz = b0z + cz - m;
        uint16_t *d = slab + ((int64_t)slot << 12) + ((gx & 15) | ((gy & 15) << 4) | ((gz & 15) << 8));
        if (item.shared)
            atomic_min_u16(d, v);
        else if (v < *d)
            *d = (uint16_t)v;
    }
}

__global__ void fill_u16_blocks_kernel(uint16_t *slab, const int32_t *__restrict__ b2s,
                                       const int32_t *__restrict__ bids, int64_t n, uint16_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n * 4096;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int s = b2s[bids[i >> 12]];
        if (s >= 2) slab[((int64_t)s << 12) + (i & 4095)] = v;
    }
}

__global__ void fill_u16_kernel(uint16_t *p, int64_t n, uint16_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

}  // namespace
}  // namespace msbg

using namespace msbg;

extern "C" {

int msbg_density_splat(msbg_grid *g, uint16_t *d_slab16, int64_t slab_slots, const float *d_pos, int64_t n,
                       float r_scan, int64_t *n_active) {
    MSBG_CHECK(g && d_slab16 && n_active, MSBG_EINVAL, "density splat: null argument");
    MSBG_CHECK(g->L == 4, MSBG_EINVAL, "density splat: 16^3-block grid required");
    MSBG_CHECK(r_scan > 0.0f && r_scan <= (float)kMaxMargin, MSBG_EINVAL, "density splat: need 0 < r_scan <= 6");
    MSBG_CHECK(n >= 0 && n < ((int64_t)1 << 31) && (n == 0 || d_pos), MSBG_EINVAL, "density splat: bad particles");
    MSBG_CHECK(g->n_blocks < ((int64_t)1 << 31), MSBG_EINVAL, "density splat: too many blocks");
    hipStream_t s = stream();
    // active blocks (active.rs:15-47) allocated in the layout and filled full
    int rc;
    int32_t *d_act = (int32_t *)workspace(10, sizeof(int32_t) * std::max<int64_t>(g->n_blocks, 1), &rc);
    if (rc) return rc;
    int64_t k = 0;
    if (n) {
        rc = msbg_active_blocks(g->sx, g->sy, g->sz, d_pos, n, r_scan, d_act, &k);
        if (rc) return rc;
    }
    std::vector<int32_t> h((size_t)k);
    if (k) MSBG_HIP(hipMemcpy(h.data(), d_act, sizeof(int32_t) * k, hipMemcpyDeviceToHost));
    std::vector<int64_t> bids(h.begin(), h.end());
    if (k) {
        rc = msbg_grid_ensure_blocks(g, bids.data(), k);
        if (rc) return rc;
    }
    MSBG_CHECK(g->n_slots <= slab_slots, MSBG_ENOMEM, "density splat: the u16 slab has fewer slots than the grid");
    *n_active = k;
    fill_u16_kernel<<<1, 256, 0, s>>>(d_slab16, 4096, 0);               // empty dummy: Density(0)
    fill_u16_kernel<<<1, 256, 0, s>>>(d_slab16 + 4096, 4096, 0xffff);  // full dummy: Density(u16::MAX)
    if (k)
        fill_u16_blocks_kernel<<<(int)std::min<int64_t>(4096, (k * 4096 + 255) / 256), 256, 0, s>>>(
            d_slab16, g->d_b2s, d_act, k, 0xffff);
    MSBG_HIP(hipGetLastError());
    if (!n) return MSBG_OK;
    // bucket the particles by block (particles/sort.rs:174-272) and gather them
    int32_t *d_bid = (int32_t *)workspace(11, sizeof(int32_t) * n, &rc);
    if (rc) return rc;
    int32_t *d_perm = (int32_t *)workspace(12, sizeof(int32_t) * n, &rc);
    if (rc) return rc;
    int32_t *d_blocks = (int32_t *)workspace(13, sizeof(int32_t) * (n + 1) * 2, &rc);
    if (rc) return rc;
    int32_t *d_starts = d_blocks + n + 1;
    float *d_sorted = (float *)workspace(14, sizeof(float) * 3 * n, &rc);
    if (rc) return rc;
    rc = msbg_particle_bids(g->sx, g->sy, g->sz, d_pos, n, d_bid);
    if (rc) return rc;
    int64_t nb = 0;
    rc = msbg_bucket_indices(d_bid, n, g->n_blocks, d_perm, d_blocks, d_starts, &nb);
    if (rc) return rc;
    rc = msbg_gather_rows(d_pos, (const uint32_t *)d_perm, n, 12, d_sorted);
    if (rc) return rc;
    SplatParams P;
    P.sx = (int)g->sx;
    P.sy = (int)g->sy;
    P.sz = (int)g->sz;
    P.nx = (int)g->nx;
    P.ny = (int)g->ny;
    P.m = (int)ceilf(r_scan);
    P.W = 16 + 2 * P.m;
    P.r = r_scan;
    P.dsq_max = r_scan * r_scan;
    P.inv = 1.0f / P.dsq_max;
    // colour lists of work items (the 8 parity buckets of splat.rs:114-123),
    // each block's particles cut into items of at most kItemParticles
    std::vector<int32_t> hb((size_t)nb), hs((size_t)nb + 1);
    MSBG_HIP(hipMemcpyAsync(hb.data(), d_blocks, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, s));
    MSBG_HIP(hipMemcpyAsync(hs.data(), d_starts, sizeof(int32_t) * (nb + 1), hipMemcpyDeviceToHost, s));
    MSBG_HIP(hipStreamSynchronize(s));
    std::vector<DItem> order;
    int64_t cnt[8] = {0};
    for (int c = 0; c < 8; ++c)
        for (int64_t i = 0; i < nb; ++i) {
            const int64_t b = hb[i];
            const int64_t bx = b % g->nx, by = (b / g->nx) % g->ny, bz = b / (g->nx * g->ny);
            if ((int)((bx & 1) | ((by & 1) << 1) | ((bz & 1) << 2)) != c) continue;
            const int32_t p0 = hs[i], p1 = hs[i + 1];
            const int32_t shared = p1 - p0 > kItemParticles ? 1 : 0;
            for (int32_t q = p0; q < p1; q += kItemParticles) {
                order.push_back(DItem{(int32_t)i, q, std::min(p1, q + kItemParticles), shared});
                ++cnt[c];
            }
        }
    MSBG_CHECK((int64_t)order.size() * kZSlabs < ((int64_t)1 << 31), MSBG_EINVAL,
               "density splat: too many work items");
    DItem *d_items = (DItem *)workspace(23, sizeof(DItem) * std::max<size_t>(order.size(), 1), &rc);
    if (rc) return rc;
    MSBG_HIP(hipMemcpyAsync(d_items, order.data(), sizeof(DItem) * order.size(), hipMemcpyHostToDevice, s));
    int64_t off = 0;
    for (int c = 0; c < 8; ++c) {
        if (cnt[c])
            density_splat_kernel<<<(int)(cnt[c] * kZSlabs), 256, 0, s>>>(P, g->d_b2s, d_slab16, d_sorted, d_blocks,
                                                                       d_items + off);
        off += cnt[c];
    }
    MSBG_HIP(hipGetLastError());
    MSBG_HIP(hipStreamSynchronize(s));
    return MSBG_OK;
}

}  // extern "C"

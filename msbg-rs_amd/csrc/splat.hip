// Density min-splat: particles::splat (particles/splat.rs:92-319) for
// D = Density (u16, channel.rs:303-375) into a u16 slab over a grid layout.
//
// Each work-group takes one block's particle bucket: it stages the block's
// (16 + 2m)^3 window (m = ceil(r_scan)) in LDS as u32 cells initialised to
// 65535, every (particle, footprint row) pair lowers the cells of its row with
// an LDS atomic min of the quantised distance, then the window is committed to
// the 3x3x3 blocks it covers with dst = min(dst, staging).  Buckets run in 8
// colour passes (block-coordinate parity), so same-colour windows never
// overlap (16 + 2m <= 32) and the commit needs no global atomics.  min is
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

constexpr int kMaxMargin = 6;  // (16 + 12)^3 u32 = 87.8 KB of LDS

__global__ __launch_bounds__(256) void density_splat_kernel(SplatParams P, const int32_t *__restrict__ b2s,
                                                            uint16_t *slab, const float *__restrict__ pos,
                                                            const int32_t *__restrict__ blocks,
                                                            const int32_t *__restrict__ starts, int color) {
    extern __shared__ uint32_t st[];
    __shared__ int rslot[27];
    const int k = blockIdx.x;
    const int bid = blocks[k];
    const int bx = bid % P.nx, by = (bid / P.nx) % P.ny, bz = bid / (P.nx * P.ny);
    if (((bx & 1) | ((by & 1) << 1) | ((bz & 1) << 2)) != color) return;  // uniform per work-group
    const int W = P.W, m = P.m, W3 = W * W * W;
    for (int i = threadIdx.x; i < W3; i += blockDim.x) st[i] = 0xffffu;
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
    __syncthreads();
    // stage (splat.rs:179-225): footprint ceil(p - r - 0.5) .. floor(p + r - 0.5),
    // clipped to the window; d2 = dx*dx + (dz*dz + dy*dy)
    const int p0 = starts[k], p1 = starts[k + 1];
    const int side = 2 * m + 1, rows = side * side;
    const int64_t items = (int64_t)(p1 - p0) * rows;
    for (int64_t i = threadIdx.x; i < items; i += blockDim.x) {
        const int j = p0 + (int)(i / rows), rr = (int)(i % rows);
        const float px = pos[3 * j], py = pos[3 * j + 1], pz = pos[3 * j + 2];
        const int iy1 = min(max((int)ceilf((py - P.r) - 0.5f), b0y - m), b0y + hi);
        const int iy2 = min(max((int)floorf((py + P.r) - 0.5f), b0y - m), b0y + hi);
        const int iz1 = min(max((int)ceilf((pz - P.r) - 0.5f), b0z - m), b0z + hi);
        const int iz2 = min(max((int)floorf((pz + P.r) - 0.5f), b0z - m), b0z + hi);
        const int iy = iy1 + rr % side, iz = iz1 + rr / side;
        if (iy > iy2 || iz > iz2) continue;
        const int ix1 = min(max((int)ceilf((px - P.r) - 0.5f), b0x - m), b0x + hi);
        const int ix2 = min(max((int)floorf((px + P.r) - 0.5f), b0x - m), b0x + hi);
        const float dz = (float)iz - (pz - 0.5f), dy = (float)iy - (py - 0.5f);
        const float dyz = dz * dz + dy * dy;
        const float px0 = px - 0.5f;
        uint32_t *row = st + ((iz - b0z + m) * W + (iy - b0y + m)) * W + (m - b0x);
        for (int ix = ix1; ix <= ix2; ++ix) {
            const float dx = (float)ix - px0;
            const float d2 = dx * dx + dyz;
            if (d2 <= P.dsq_max) atomicMin(row + ix, (uint32_t)(int)roundf((d2 * P.inv) * 65535.0f));
        }
    }
    __syncthreads();
    // commit (splat.rs:227-313): each touched window cell, dst = min(dst, staged)
    for (int c = threadIdx.x; c < W3; c += blockDim.x) {
        const uint32_t v = st[c];
        if (v == 0xffffu) continue;
        const int cx = c % W, cy = (c / W) % W, cz = c / (W * W);
        const int rx = cx < m ? 0 : (cx < m + 16 ? 1 : 2), ry = cy < m ? 0 : (cy < m + 16 ? 1 : 2),
                  rz = cz < m ? 0 : (cz < m + 16 ? 1 : 2);
        const int slot = rslot[rx + 3 * ry + 9 * rz];
        if (slot < 0) continue;
        const int gx = b0x + cx - m, gy = b0y + cy - m, gz = b0z + cz - m;
        uint16_t *d = slab + ((int64_t)slot << 12) + ((gx & 15) | ((gy & 15) << 4) | ((gz & 15) << 8));
        if (v < *d) *d = (uint16_t)v;
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
    const size_t lds = sizeof(uint32_t) * P.W * P.W * P.W;
    if (lds > 65536)
        MSBG_HIP(hipFuncSetAttribute((const void *)density_splat_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds));
    for (int color = 0; color < 8 && nb; ++color)
        density_splat_kernel<<<(int)nb, 256, lds, s>>>(P, g->d_b2s, d_slab16, d_sorted, d_blocks, d_starts, color);
    MSBG_HIP(hipGetLastError());
    MSBG_HIP(hipStreamSynchronize(s));
    return MSBG_OK;
}

}  // extern "C"

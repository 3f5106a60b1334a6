// Multires level transfers: 2x2x2 restriction, the 2:1 level-map
// regularization and the per-block topology flags.
#include <algorithm>
#include <vector>

#include "internal.hpp"
#include "primitives.hpp"

namespace msbg {

// downsample_channel_avg -- multires/downsample.rs:33-57.  Workgroup k owns
// coarse block bids[k]; each coarse voxel sums its 8 children in dz, dy, dx
// order and divides by 8.
template <int LF>
__global__ __launch_bounds__(256) void downsample_kernel(GridView f, GridView c, const int64_t *bids) {
    constexpr int BC = 1 << (LF - 1), NC = BC * BC * BC, NF = 1 << (3 * LF);
    const int64_t bid = bids[blockIdx.x];
    const float *sp = f.slab + (int64_t)f.b2s[bid] * NF;
    float *dp = c.slab + (int64_t)c.b2s[bid] * NC;
    for (int v = threadIdx.x; v < NC; v += blockDim.x) {
        const int cx = v & (BC - 1), cy = (v >> (LF - 1)) & (BC - 1), cz = v >> (2 * (LF - 1));
        float sum = 0.0f;
#pragma unroll
        for (int dz = 0; dz < 2; ++dz)
#pragma unroll
            for (int dy = 0; dy < 2; ++dy)
#pragma unroll
                for (int dx = 0; dx < 2; ++dx)
                    sum += sp[(2 * cx + dx) | ((2 * cy + dy) << LF) | ((2 * cz + dz) << (2 * LF))];
        dp[v] = sum / 8.0f;
    }
}

// One relaxation pass of regularize_refinement_map (refinement.rs:72-103).
__global__ void regularize_kernel(const uint8_t *prev, uint8_t *out, int64_t nx, int64_t ny, int64_t nz, int ll,
                                  int n_levels) {
    const int64_t nxy = nx * ny, nb = nxy * nz;
    for (int64_t bid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; bid < nb;
         bid += (int64_t)gridDim.x * blockDim.x) {
        const uint8_t level = prev[bid];
        uint8_t lo = level;
        if (level >= ll) {
            const int64_t bx = bid % nx, by = (bid / nx) % ny, bz = bid / nxy;
            uint8_t level2 = (uint8_t)(n_levels - 1);
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        if (!dx && !dy && !dz) continue;
                        const int64_t x2 = bx + dx, y2 = by + dy, z2 = bz + dz;
                        if (x2 < 0 || x2 >= nx || y2 < 0 || y2 >= ny || z2 < 0 || z2 >= nz) continue;
                        const uint8_t l2 = prev[x2 + y2 * nx + z2 * nxy];
                        level2 = l2 < level2 ? l2 : level2;
                    }
            if (level2 == ll - 1) lo = (uint8_t)ll;
        }
        out[bid] = lo;
    }
}

}  // namespace msbg

using namespace msbg;

namespace msbg {

// compute_block_topology -- multires/refinement.rs:137-183, one thread per
// block: user flags | DOM_BORDER on the block-grid faces | FINE_COARSE if a
// 26-neighbour is coarser | COARSE_FINE if one is finer
// (blockinfo.rs:39-54 bit values).  exists[b] = 1 for EXISTS blocks.
constexpr uint16_t kBlkDomBorder = 1u << 1, kBlkCoarseFine = 1u << 2, kBlkFineCoarse = 1u << 3,
                   kBlkExists = 1u << 12;

__global__ __launch_bounds__(256) void block_topology_kernel(const uint8_t *levels, int64_t nx, int64_t ny,
                                                             int64_t nz, int n_levels, const uint16_t *user,
                                                             uint16_t *flags, uint32_t *exists) {
    const int64_t nxy = nx * ny, nb = nxy * nz;
    for (int64_t bid = (int64_t)blockIdx.x * 256 + threadIdx.x; bid < nb; bid += (int64_t)gridDim.x * 256) {
        const int level = levels[bid];
        uint16_t f = user ? user[bid] : (uint16_t)0;
        const int64_t bx = bid % nx, by = (bid / nx) % ny, bz = bid / nxy;
        if (bx == 0 || bx == nx - 1 || by == 0 || by == ny - 1 || bz == 0 || bz == nz - 1) f |= kBlkDomBorder;
        int lmax = 0, lmin = n_levels - 1;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int64_t x = bx + dx, y = by + dy, z = bz + dz;
                    if ((dx | dy | dz) == 0 || x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) continue;
                    const int l2 = levels[x + y * nx + z * nxy];
                    lmax = max(lmax, l2);
                    lmin = min(lmin, l2);
                }
        if (lmax > level) f |= kBlkFineCoarse;
        if (lmin < level) f |= kBlkCoarseFine;
        flags[bid] = f;
        exists[bid] = (f & kBlkExists) ? 1u : 0u;
    }
}

// mark[b] = 1 for the FINE_COARSE blocks of level l.
__global__ __launch_bounds__(256) void fc_mark_kernel(const uint16_t *flags, const uint8_t *levels, int64_t nb,
                                                      int l, uint32_t *mark) {
    for (int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x; b < nb; b += (int64_t)gridDim.x * 256)
        mark[b] = ((flags[b] & kBlkFineCoarse) && levels[b] == l) ? 1u : 0u;
}

// out[pos[b]] = b for the marked blocks (pos: exclusive scan of the marks), in
// ascending block order.
__global__ __launch_bounds__(256) void fc_scatter_kernel(const uint16_t *flags, const uint8_t *levels,
                                                         const uint32_t *pos, int64_t nb, int l, int64_t *out) {
    for (int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x; b < nb; b += (int64_t)gridDim.x * 256)
        if ((flags[b] & kBlkFineCoarse) && levels[b] == l) out[pos[b]] = b;
}

}  // namespace msbg

extern "C" {

int msbg_downsample_avg(msbg_grid *coarse, const msbg_grid *fine, const int64_t *bids, int64_t n) {
    MSBG_CHECK(coarse && fine, MSBG_EINVAL, "downsample: null level");
    MSBG_CHECK(coarse->L == fine->L - 1 && coarse->nx == fine->nx && coarse->ny == fine->ny && coarse->nz == fine->nz,
               MSBG_EINVAL, "downsample: coarse block edge must be half the fine, same block grid");
    MSBG_CHECK(n >= 0 && (n == 0 || bids), MSBG_EINVAL, "downsample: bad bid list");
    if (n == 0) return MSBG_OK;
    int rc = msbg_grid_ensure_blocks(coarse, bids, n);  // downsample.rs:26-31
    if (rc) return rc;
    int64_t *d_bids = nullptr;
    MSBG_HIP(hipMalloc(&d_bids, sizeof(int64_t) * n));
    hipError_t e = hipMemcpyAsync(d_bids, bids, sizeof(int64_t) * n, hipMemcpyHostToDevice, stream());
    if (e == hipSuccess) {
        GridView f = fine->view(), c = coarse->view();
        dim3 grid((unsigned)n), block(256);
        switch (fine->L) {
            case 4: hipLaunchKernelGGL(downsample_kernel<4>, grid, block, 0, stream(), f, c, d_bids); break;
            case 3: hipLaunchKernelGGL(downsample_kernel<3>, grid, block, 0, stream(), f, c, d_bids); break;
            case 2: hipLaunchKernelGGL(downsample_kernel<2>, grid, block, 0, stream(), f, c, d_bids); break;
            default: hipLaunchKernelGGL(downsample_kernel<1>, grid, block, 0, stream(), f, c, d_bids); break;
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream());
    (void)hipFree(d_bids);
    MSBG_HIP(e);
    return MSBG_OK;
}

int msbg_regularize_levels(uint8_t *dev_levels, int64_t nx, int64_t ny, int64_t nz, int n_levels) {
    MSBG_CHECK(dev_levels && nx > 0 && ny > 0 && nz > 0, MSBG_EINVAL, "regularize: bad arguments");
    MSBG_CHECK(n_levels >= 2 && n_levels <= 5, MSBG_EINVAL, "regularize: n_levels must be 2..5 (grid.rs:45)");
    if (n_levels <= 2) return MSBG_OK;
    const int64_t nb = nx * ny * nz;
    uint8_t *prev = nullptr;
    MSBG_HIP(hipMalloc(&prev, nb));
    hipError_t e = hipMemcpyAsync(prev, dev_levels, nb, hipMemcpyDeviceToDevice, stream());
    const unsigned blocks = (unsigned)std::min<int64_t>((nb + 255) / 256, 65535);
    for (int ll = 1; ll <= n_levels - 2 && e == hipSuccess; ++ll) {
        hipLaunchKernelGGL(regularize_kernel, dim3(blocks), dim3(256), 0, stream(), prev, dev_levels, nx, ny, nz, ll,
                           n_levels);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(prev, dev_levels, nb, hipMemcpyDeviceToDevice, stream());
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream());
    (void)hipFree(prev);
    MSBG_HIP(e);
    return MSBG_OK;
}

int msbg_block_topology(const uint8_t *d_levels, int64_t nx, int64_t ny, int64_t nz, int n_levels,
                        const uint16_t *d_user_flags, uint16_t *d_flags_out, int64_t *d_fc_bids_out,
                        int64_t *fc_counts, int64_t *n_act) {
    MSBG_CHECK(d_levels && d_flags_out && d_fc_bids_out && fc_counts && n_act, MSBG_EINVAL,
               "block topology: null argument");
    MSBG_CHECK(nx > 0 && ny > 0 && nz > 0 && nx * ny * nz < ((int64_t)1 << 31), MSBG_EINVAL,
               "block topology: bad block grid");
    MSBG_CHECK(n_levels >= 1 && n_levels <= 8, MSBG_EINVAL, "block topology: n_levels must be 1..8");
    const int64_t nb = nx * ny * nz;
    int rc;
    uint32_t *mark = (uint32_t *)workspace(18, sizeof(uint32_t) * (size_t)nb, &rc);
    if (rc) return rc;
    const unsigned grid = (unsigned)std::min<int64_t>((nb + 255) / 256, 65535);
    hipStream_t s = stream();
    hipLaunchKernelGGL(block_topology_kernel, dim3(grid), dim3(256), 0, s, d_levels, nx, ny, nz, n_levels, d_user_flags,
                       d_flags_out, mark);
    MSBG_HIP(hipGetLastError());
    uint32_t total = 0;
    rc = scan_exclusive_u32(mark, mark, nb, &total);
    if (rc) return rc;
    *n_act = total;
    int64_t off = 0;
    for (int l = 0; l < n_levels; ++l) {
        hipLaunchKernelGGL(fc_mark_kernel, dim3(grid), dim3(256), 0, s, d_flags_out, d_levels, nb, l, mark);
        MSBG_HIP(hipGetLastError());
        rc = scan_exclusive_u32(mark, mark, nb, &total);
        if (rc) return rc;
        hipLaunchKernelGGL(fc_scatter_kernel, dim3(grid), dim3(256), 0, s, d_flags_out, d_levels, mark, nb, l,
                           d_fc_bids_out + off);
        MSBG_HIP(hipGetLastError());
        fc_counts[l] = total;
        off += total;
    }
    MSBG_HIP(hipStreamSynchronize(s));
    return MSBG_OK;
}

}  // extern "C"

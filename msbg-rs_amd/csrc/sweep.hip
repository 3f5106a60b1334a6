// Halo gather and the 7/19/25-tap block smoothers (8-colour, in place).
//
// One workgroup owns one 16^3 block: it stages the block plus its halo
// (HSX^3 floats, HSX = 16 + 2*HALO, 23.3 KB or 32 KB) in LDS, resolving the
// (2H+1)^3 neighbour blocks once through the bid->slot table, then 256 lanes
// each update a 16-cell z column and write it back in place.  Blocks of one
// colour never share a halo (solver.rs:242-246), so a colour is one launch
// and launches on one stream order the colour passes.
//
// The arithmetic repeats the reference's operation order exactly (fma where
// it writes mul_add, no contraction elsewhere; the file is built with
// -ffp-contract=off), so results are bit-identical to it.
#include <algorithm>
#include <cstring>

#include "internal.hpp"

namespace msbg {

constexpr int kThreads = 256;

// Resolves the (2H+1)^3 neighbourhood of block (bx0,by0,bz0) into LDS:
// s_nb = fine slot (0 = empty dummy outside the domain, halo.rs:59-80) and,
// for multires, s_nbc = coarse slot of a coarser neighbour or -1
// (halo.rs:234-251).
template <int HALO, bool MULTI>
__device__ __forceinline__ void resolve_neighbours(const GridView &g, const GridView &c, const uint8_t *levels,
                                                   int bx0, int by0, int bz0, int64_t bid, int *s_nb,
                                                   int *s_nbc) {
    constexpr int n = 2 * HALO + 1;
    const int t = threadIdx.x;
    if (t < n * n * n) {
        int dx = t % n, dy = (t / n) % n, dz = t / (n * n);
        int bx = bx0 + dx - HALO, by = by0 + dy - HALO, bz = bz0 + dz - HALO;
        int slot = 0, cslot = -1;
        if (bx >= 0 && bx < g.nx && by >= 0 && by < g.ny && bz >= 0 && bz < g.nz) {
            int64_t b = (int64_t)bx + (int64_t)by * g.nx + (int64_t)bz * g.nx * g.ny;
            if (MULTI && levels[b] > levels[bid])
                cslot = c.b2s[b];
            else
                slot = g.b2s[b];
        }
        s_nb[t] = slot;
        if (MULTI) s_nbc[t] = cslot;
    }
}

// Fills the LDS tile (halo.rs:31-170 / halo.rs:187-358).  `skip_corners`
// leaves cells with two or more halo axes unread (a faces-only 7-point
// stencil never reads them); `zero_faces_only_holes` writes 0 where a
// faces-only fill leaves a fresh HaloBlock untouched (halo.rs:154).
template <int HALO, bool MULTI>
__device__ __forceinline__ void fill_tile(float *tile, const GridView &g, const GridView &c, const int *s_nb,
                                          const int *s_nbc, int bx0, int by0, int bz0, int bc, float empty_f32,
                                          bool full, bool skip_corners) {
    constexpr int B = 16, L = 4, HSX = B + 2 * HALO, n = 2 * HALO + 1;
    constexpr int NT = HSX * HSX * HSX;
    const int x0 = bx0 * B, y0 = by0 * B, z0 = bz0 * B;
    const int smax = g.sx - 1, symax = g.sy - 1, szmax = g.sz - 1;
    for (int i = threadIdx.x; i < NT; i += kThreads) {
        const int hx = i % HSX, hy = (i / HSX) % HSX, hz = i / (HSX * HSX);
        const bool xh = hx < HALO || hx >= HALO + B;
        const bool yh = hy < HALO || hy >= HALO + B;
        const bool zh = hz < HALO || hz >= HALO + B;
        if (skip_corners && ((int)xh + (int)yh + (int)zh) >= 2) continue;
        if (!full && xh && (yh || zh)) {
            tile[i] = 0.0f;
            continue;
        }
        int rx, ry, rz;
        float v = empty_f32;
        if (resolve_axis(x0 + hx - HALO, smax, bc, rx) && resolve_axis(y0 + hy - HALO, symax, bc, ry) &&
            resolve_axis(z0 + hz - HALO, szmax, bc, rz)) {
            const int nb = ((rx >> L) - bx0 + HALO) + n * ((ry >> L) - by0 + HALO) + n * n * ((rz >> L) - bz0 + HALO);
            if (MULTI && s_nbc[nb] >= 0) {
                // Coarse ghost: fine coordinate >> 1 inside the 8^3 coarse block.
                const int cvid = ((rx >> 1) & 7) | (((ry >> 1) & 7) << 3) | (((rz >> 1) & 7) << 6);
                v = c.slab[(int64_t)s_nbc[nb] * 512 + cvid];
            } else {
                const int vid = (rx & 15) | ((ry & 15) << 4) | ((rz & 15) << 8);
                v = g.slab[(int64_t)s_nb[nb] * 4096 + vid];
            }
        }
        tile[i] = v;
    }
}

__device__ __forceinline__ float clamp_f(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

// kernel_laplacian -- math/laplacian.rs:39-69.
__device__ __forceinline__ float lap7(const float *t, int h, float dt, bool clamp01) {
    constexpr int dy = 18, dz = 18 * 18;
    const float c = t[h], lx = t[h - 1], rx = t[h + 1], dyu = t[h - dy], dyd = t[h + dy], dzb = t[h - dz],
                dzf = t[h + dz];
    const float nb = lx + rx + dyu + dyd + dzb + dzf;
    const float lap = __builtin_fmaf(c, -6.0f, nb);
    const float fx = (rx - lx) * 0.5f, fy = (dyd - dyu) * 0.5f, fz = (dzf - dzb) * 0.5f;
    const float g = __builtin_sqrtf(__builtin_fmaf(fx, fx, __builtin_fmaf(fy, fy, fz * fz)));
    float v = __builtin_fmaf(dt * lap, g, c);
    return clamp01 ? clamp_f(v, 0.0f, 1.0f) : v;
}

// kernel_meancurv -- math/meancurv.rs:49-100.
__device__ __forceinline__ float mc19(const float *t, int h, float dt, bool clamp01) {
    constexpr int dy = 18, dz = 18 * 18;
    const float f0 = t[h], f1 = t[h - 1], f2 = t[h + 1], f3 = t[h - dy], f4 = t[h + dy], f5 = t[h - dz],
                f6 = t[h + dz];
    const float fx = (f2 - f1) * 0.5f, fy = (f4 - f3) * 0.5f, fz = (f6 - f5) * 0.5f;
    const float fxx = f1 + f2 - 2.0f * f0, fyy = f3 + f4 - 2.0f * f0, fzz = f5 + f6 - 2.0f * f0;
    const float d1 = t[h + 1 + dy], d2 = t[h + 1 - dy], d3 = t[h - 1 + dy], d4 = t[h - 1 - dy];
    const float d5 = t[h + 1 + dz], d6 = t[h + 1 - dz], d7 = t[h - 1 + dz], d8 = t[h - 1 - dz];
    const float d9 = t[h + dy + dz], d10 = t[h + dy - dz], d11 = t[h - dy + dz], d12 = t[h - dy - dz];
    const float fyx = (d1 - d2 - d3 + d4) * 0.25f;
    const float fzx = (d5 - d6 - d7 + d8) * 0.25f;
    const float fzy = (d9 - d10 - d11 + d12) * 0.25f;
    const float cross = fx * (fy * fyx) + fx * (fz * fzx) + fy * (fz * fzy);
    const float hnum = __builtin_fmaf(fy * fy + fz * fz, fxx,
                                      __builtin_fmaf(fx * fx + fz * fz, fyy,
                                                     __builtin_fmaf(fx * fx + fy * fy, fzz, -2.0f * cross)));
    const float grad = __builtin_fmaf(fx, fx, __builtin_fmaf(fy, fy, fz * fz));
    const float hc = grad > 1e-7f ? hnum / grad : 0.0f;
    const float d = clamp_f(dt * hc, -0.1f, 0.1f);
    const float v = f0 + d;
    return clamp01 ? clamp_f(v, 0.0f, 1.0f) : v;
}

// kernel_bilaplacian -- math/bilaplacian.rs:47-88.
__device__ __forceinline__ float bl25(const float *t, int h, float dt, bool clamp01) {
    constexpr int dy = 20, dz = 20 * 20;
    const float f0 = t[h], f1 = t[h - 1], f2 = t[h + 1], f3 = t[h - dy], f4 = t[h + dy], f5 = t[h - dz],
                f6 = t[h + dz];
    const float s1 = f1 + f2 + f3 + f4 + f5 + f6;
    const float s2 = t[h - 2] + t[h + 2] + t[h - 2 * dy] + t[h + 2 * dy] + t[h - 2 * dz] + t[h + 2 * dz];
    const float sd = t[h + 1 + dy] + t[h + 1 - dy] + t[h - 1 + dy] + t[h - 1 - dy] + t[h + 1 + dz] + t[h + 1 - dz] +
                     t[h - 1 + dz] + t[h - 1 - dz] + t[h + dy + dz] + t[h + dy - dz] + t[h - dy + dz] +
                     t[h - dy - dz];
    const float hh = 42.0f * f0 - 12.0f * s1 + s2 + 2.0f * sd;
    const float fx = (f2 - f1) * 0.5f, fy = (f4 - f3) * 0.5f, fz = (f6 - f5) * 0.5f;
    const float g = __builtin_sqrtf(__builtin_fmaf(fx, fx, __builtin_fmaf(fy, fy, fz * fz)));
    const float upd = (dt * hh) * g;
    const float d = clamp_f(-upd, -0.1f, 0.1f);
    const float v = f0 + d;
    return clamp01 ? clamp_f(v, 0.0f, 1.0f) : v;
}

// One colour pass: workgroup k updates block bids[k] in place.
template <int STENCIL, bool MULTI>
__global__ __launch_bounds__(kThreads) void sweep_kernel(GridView g, GridView c, const uint8_t *levels,
                                                         const int64_t *bids, float dt, int clamp01,
                                                         float empty_f32) {
    constexpr int HALO = STENCIL == MSBG_BILAPLACIAN ? 2 : 1;
    constexpr int HSX = 16 + 2 * HALO;
    __shared__ float tile[HSX * HSX * HSX];
    __shared__ int s_nb[125];
    __shared__ int s_nbc[MULTI ? 125 : 1];
    const int64_t bid = bids[blockIdx.x];
    const int slot = g.b2s[bid];
    if (slot < 2) return;  // dummy or absent: solver.rs:201-204 (uniform per workgroup)
    const int bx0 = (int)(bid % g.nx), by0 = (int)((bid / g.nx) % g.ny), bz0 = (int)(bid / ((int64_t)g.nx * g.ny));
    resolve_neighbours<HALO, MULTI>(g, c, levels, bx0, by0, bz0, bid, s_nb, s_nbc);
    __syncthreads();
    fill_tile<HALO, MULTI>(tile, g, c, s_nb, s_nbc, bx0, by0, bz0, MSBG_BC_NEUMANN, empty_f32, true,
                           STENCIL == MSBG_LAPLACIAN);
    __syncthreads();
    const int x = threadIdx.x & 15, y = threadIdx.x >> 4;
    float *out = g.slab + (int64_t)slot * 4096 + y * 16 + x;
    const bool cl = clamp01 != 0;
#pragma unroll 4
    for (int z = 0; z < 16; ++z) {
        const int h = ((z + HALO) * HSX + (y + HALO)) * HSX + (x + HALO);
        float v;
        if (STENCIL == MSBG_LAPLACIAN)
            v = lap7(tile, h, dt, cl);
        else if (STENCIL == MSBG_MEAN_CURVATURE)
            v = mc19(tile, h, dt, cl);
        else
            v = bl25(tile, h, dt, cl);
        out[z * 256] = v;
    }
}

// Standalone halo gather: workgroup k writes the tile of bids[k].
template <int HALO, bool MULTI>
__global__ __launch_bounds__(kThreads) void gather_kernel(GridView g, GridView c, const uint8_t *levels,
                                                          const int64_t *bids, int bc, int full, float empty_f32,
                                                          float *tiles) {
    constexpr int HSX = 16 + 2 * HALO;
    constexpr int NT = HSX * HSX * HSX;
    __shared__ float tile[NT];
    __shared__ int s_nb[125];
    __shared__ int s_nbc[MULTI ? 125 : 1];
    const int64_t bid = bids[blockIdx.x];
    const int bx0 = (int)(bid % g.nx), by0 = (int)((bid / g.nx) % g.ny), bz0 = (int)(bid / ((int64_t)g.nx * g.ny));
    resolve_neighbours<HALO, MULTI>(g, c, levels, bx0, by0, bz0, bid, s_nb, s_nbc);
    __syncthreads();
    fill_tile<HALO, MULTI>(tile, g, c, s_nb, s_nbc, bx0, by0, bz0, bc, empty_f32, full != 0, false);
    __syncthreads();
    float *dst = tiles + (int64_t)blockIdx.x * NT;
    for (int i = threadIdx.x; i < NT; i += kThreads) dst[i] = tile[i];
}

}  // namespace msbg

using namespace msbg;

struct msbg_sweep_plan {
    msbg_grid *g = nullptr;
    const msbg_grid *coarse = nullptr;
    const uint8_t *d_levels = nullptr;
    int64_t *d_bids = nullptr;
    int64_t counts[8] = {0};
    int64_t n = 0;
};

namespace {

uint64_t spread21(uint64_t v) {  // multires/sort.rs:8-16
    v &= 0x1fffffULL;
    v = (v | (v << 32)) & 0x001f00000000ffffULL;
    v = (v | (v << 16)) & 0x001f0000ff0000ffULL;
    v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
    v = (v | (v << 2)) & 0x1249249249249249ULL;
    return v;
}

int color_buckets(const msbg_grid *g, const int64_t *active, int64_t n, int64_t *out, int64_t *counts) {
    MSBG_CHECK(g && counts && (n == 0 || (active && out)), MSBG_EINVAL, "color_buckets: null argument");
    std::vector<std::pair<uint64_t, int64_t>> b[8];
    for (int64_t i = 0; i < n; ++i) {
        int64_t bid = active[i];
        MSBG_CHECK(bid >= 0 && bid < g->n_blocks, MSBG_EINVAL, "color_buckets: bid out of range");
        uint64_t bx = bid % g->nx, by = (bid / g->nx) % g->ny, bz = bid / g->nxy;
        int color = (int)((bx & 1) | ((by & 1) << 1) | ((bz & 1) << 2));
        b[color].push_back({spread21(bx) | (spread21(by) << 1) | (spread21(bz) << 2), bid});
    }
    int64_t k = 0;
    for (int c = 0; c < 8; ++c) {
        std::sort(b[c].begin(), b[c].end());
        counts[c] = (int64_t)b[c].size();
        for (auto &kv : b[c]) out[k++] = kv.second;
    }
    return MSBG_OK;
}

int check_stencil(int stencil) {
    MSBG_CHECK(stencil >= 0 && stencil <= 2, MSBG_EINVAL, "unknown stencil");
    return MSBG_OK;
}

int plan_new(msbg_grid *g, const msbg_grid *coarse, const uint8_t *d_levels, const int64_t *active, int64_t n,
             msbg_sweep_plan **out) {
    MSBG_CHECK(out, MSBG_EINVAL, "null out");
    *out = nullptr;
    MSBG_CHECK(g && g->L == 4, MSBG_EINVAL, "sweeps run on 16^3-block grids");
    if (coarse) {
        MSBG_CHECK(coarse->L == 3 && coarse->nx == g->nx && coarse->ny == g->ny && coarse->nz == g->nz, MSBG_EINVAL,
                   "multires sweep: coarse level must be 8^3 blocks over the same block grid");
        MSBG_CHECK(d_levels, MSBG_EINVAL, "multires sweep: null level map");
    }
    MSBG_CHECK(n >= 0, MSBG_EINVAL, "negative active count");
    std::vector<int64_t> order((size_t)std::max<int64_t>(n, 1));
    msbg_sweep_plan *p = new msbg_sweep_plan();
    int rc = color_buckets(g, active, n, order.data(), p->counts);
    if (rc) {
        delete p;
        return rc;
    }
    p->g = g;
    p->coarse = coarse;
    p->d_levels = d_levels;
    p->n = n;
    if (n > 0) {
        hipError_t e = hipMalloc(&p->d_bids, sizeof(int64_t) * n);
        if (e == hipSuccess)
            e = hipMemcpyAsync(p->d_bids, order.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, stream());
        if (e == hipSuccess) e = hipStreamSynchronize(stream());
        if (e != hipSuccess) {
            msbg_sweep_plan_free(p);
            return hip_fail(e, "sweep plan upload");
        }
    }
    *out = p;
    return MSBG_OK;
}

template <bool MULTI>
int launch_sweep(msbg_sweep_plan *p, int stencil, float dt, int clamp01) {
    GridView g = p->g->view();
    GridView c = MULTI ? p->coarse->view() : g;
    const float empty_f32 = MULTI ? 0.0f : p->g->empty_value;
    int64_t off = 0;
    for (int col = 0; col < 8; ++col) {
        const int64_t cnt = p->counts[col];
        if (cnt > 0) {
            dim3 grid((unsigned)cnt), block(kThreads);
            const int64_t *bids = p->d_bids + off;
            switch (stencil) {
                case MSBG_LAPLACIAN:
                    hipLaunchKernelGGL((sweep_kernel<MSBG_LAPLACIAN, MULTI>), grid, block, 0, stream(), g, c,
                                       p->d_levels, bids, dt, clamp01, empty_f32);
                    break;
                case MSBG_MEAN_CURVATURE:
                    hipLaunchKernelGGL((sweep_kernel<MSBG_MEAN_CURVATURE, MULTI>), grid, block, 0, stream(), g, c,
                                       p->d_levels, bids, dt, clamp01, empty_f32);
                    break;
                default:
                    hipLaunchKernelGGL((sweep_kernel<MSBG_BILAPLACIAN, MULTI>), grid, block, 0, stream(), g, c,
                                       p->d_levels, bids, dt, clamp01, empty_f32);
            }
            MSBG_HIP(hipGetLastError());
        }
        off += cnt;
    }
    return MSBG_OK;
}

int gather(const msbg_grid *g, const msbg_grid *coarse, const uint8_t *d_levels, const int64_t *bids, int64_t n,
           int halo, int full, int bc, float *d_tiles) {
    MSBG_CHECK(g && g->L == 4, MSBG_EINVAL, "halo gather runs on 16^3-block grids");
    MSBG_CHECK(halo == 1 || halo == 2, MSBG_EINVAL, "halo must be 1 or 2 (halo.rs:62)");
    MSBG_CHECK(bc >= 0 && bc <= 2, MSBG_EINVAL, "unknown boundary condition");
    MSBG_CHECK(n >= 0 && (n == 0 || (bids && d_tiles)), MSBG_EINVAL, "halo gather: null argument");
    if (coarse)
        MSBG_CHECK(coarse->L == 3 && coarse->nx == g->nx && coarse->ny == g->ny && coarse->nz == g->nz && d_levels,
                   MSBG_EINVAL, "multires gather: coarse level must be 8^3 blocks over the same block grid");
    for (int64_t i = 0; i < n; ++i)
        MSBG_CHECK(bids[i] >= 0 && bids[i] < g->n_blocks, MSBG_EINVAL, "halo gather: bid out of range");
    if (n == 0) return MSBG_OK;
    int64_t *d_bids = nullptr;
    MSBG_HIP(hipMalloc(&d_bids, sizeof(int64_t) * n));
    hipError_t e = hipMemcpyAsync(d_bids, bids, sizeof(int64_t) * n, hipMemcpyHostToDevice, stream());
    if (e == hipSuccess) {
        GridView gv = g->view();
        GridView cv = coarse ? coarse->view() : gv;
        const float empty_f32 = coarse ? 0.0f : g->empty_value;
        dim3 grid((unsigned)n), block(kThreads);
        if (coarse) {
            if (halo == 1)
                hipLaunchKernelGGL((gather_kernel<1, true>), grid, block, 0, stream(), gv, cv, d_levels, d_bids, bc,
                                   full, empty_f32, d_tiles);
            else
                hipLaunchKernelGGL((gather_kernel<2, true>), grid, block, 0, stream(), gv, cv, d_levels, d_bids, bc,
                                   full, empty_f32, d_tiles);
        } else {
            if (halo == 1)
                hipLaunchKernelGGL((gather_kernel<1, false>), grid, block, 0, stream(), gv, cv, d_levels, d_bids,
                                   bc, full, empty_f32, d_tiles);
            else
                hipLaunchKernelGGL((gather_kernel<2, false>), grid, block, 0, stream(), gv, cv, d_levels, d_bids,
                                   bc, full, empty_f32, d_tiles);
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream());
    hipFree(d_bids);
    MSBG_HIP(e);
    return MSBG_OK;
}

}  // namespace

extern "C" {

int msbg_color_buckets(const msbg_grid *g, const int64_t *active, int64_t n, int64_t *out, int64_t *counts) {
    return color_buckets(g, active, n, out, counts);
}

int msbg_sweep_plan_new(msbg_grid *g, const int64_t *active, int64_t n, msbg_sweep_plan **out) {
    return plan_new(g, nullptr, nullptr, active, n, out);
}

int msbg_sweep_plan_new_multires(msbg_grid *fine, const msbg_grid *coarse, const uint8_t *dev_levels,
                                 const int64_t *active, int64_t n, msbg_sweep_plan **out) {
    MSBG_CHECK(coarse, MSBG_EINVAL, "multires plan: null coarse level");
    return plan_new(fine, coarse, dev_levels, active, n, out);
}

int msbg_sweep_plan_free(msbg_sweep_plan *p) {
    if (!p) return MSBG_OK;
    if (p->d_bids) hipFree(p->d_bids);
    delete p;
    return MSBG_OK;
}

int msbg_sweep_plan_run_async(msbg_sweep_plan *p, int stencil, float dt, int iterations, int clamp01) {
    MSBG_CHECK(p, MSBG_EINVAL, "null plan");
    int rc = check_stencil(stencil);
    if (rc) return rc;
    MSBG_CHECK(iterations >= 0, MSBG_EINVAL, "negative iterations");
    for (int it = 0; it < iterations; ++it) {
        rc = p->coarse ? launch_sweep<true>(p, stencil, dt, clamp01) : launch_sweep<false>(p, stencil, dt, clamp01);
        if (rc) return rc;
    }
    return MSBG_OK;
}

int msbg_sweep(msbg_grid *g, const int64_t *active, int64_t n, int stencil, float dt, int iterations, int clamp01) {
    int rc = check_stencil(stencil);
    if (rc) return rc;
    msbg_sweep_plan *p = nullptr;
    rc = plan_new(g, nullptr, nullptr, active, n, &p);
    if (rc) return rc;
    rc = msbg_sweep_plan_run_async(p, stencil, dt, iterations, clamp01);
    if (rc == MSBG_OK) {
        hipError_t e = hipStreamSynchronize(stream());
        if (e != hipSuccess) rc = hip_fail(e, "msbg_sweep");
    }
    msbg_sweep_plan_free(p);
    return rc;
}

int msbg_sweep_multires(msbg_grid *fine, const msbg_grid *coarse, const uint8_t *dev_levels, const int64_t *active,
                        int64_t n, int stencil, float dt, int iterations, int clamp01) {
    int rc = check_stencil(stencil);
    if (rc) return rc;
    msbg_sweep_plan *p = nullptr;
    rc = msbg_sweep_plan_new_multires(fine, coarse, dev_levels, active, n, &p);
    if (rc) return rc;
    rc = msbg_sweep_plan_run_async(p, stencil, dt, iterations, clamp01);
    if (rc == MSBG_OK) {
        hipError_t e = hipStreamSynchronize(stream());
        if (e != hipSuccess) rc = hip_fail(e, "msbg_sweep_multires");
    }
    msbg_sweep_plan_free(p);
    return rc;
}

int msbg_halo_gather(const msbg_grid *g, const int64_t *bids, int64_t n, int halo, int full, int bc,
                     float *dev_tiles) {
    return gather(g, nullptr, nullptr, bids, n, halo, full, bc, dev_tiles);
}

int msbg_halo_gather_multires(const msbg_grid *fine, const msbg_grid *coarse, const uint8_t *dev_levels,
                              const int64_t *bids, int64_t n, int halo, int full, int bc, float *dev_tiles) {
    MSBG_CHECK(coarse, MSBG_EINVAL, "multires gather: null coarse level");
    return gather(fine, coarse, dev_levels, bids, n, halo, full, bc, dev_tiles);
}

}  // extern "C"

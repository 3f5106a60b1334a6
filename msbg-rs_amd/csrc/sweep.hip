// Halo gather and the 7/19/25-tap block smoothers (8-colour, in place).
//
// One workgroup owns one 16^3 block: it stages the block plus its halo in
// LDS and updates the block in place.  Blocks of one colour never share a
// halo (solver.rs:242-246), so a colour is one launch, and launches on one
// stream order the colour passes.
//
// A sweep plan resolves, once, the halo topology of every planned block: the
// slab slot of each of its (2H+1)^3 neighbours (halo.rs:59-80), or the coarse
// slot of a coarser neighbour (halo.rs:234-251).  A workgroup then reads one
// table row instead of chasing bid -> slot -> level pointers, and the colour
// passes of a sweep replay as one HIP graph.
//
// The arithmetic repeats the reference's operation order exactly (fma where
// it writes mul_add, no contraction elsewhere; the file is built with
// -ffp-contract=off), so results are bit-identical to it.
#include <algorithm>
#include <cstring>
#include <vector>

#include "internal.hpp"
#include "primitives.hpp"

namespace msbg {

constexpr int kGatherThreads = 256;
constexpr int kSweepThreads = 512;
// LDS row pitch of the sweep tile: the 16-float block row sits at [4, 20) so
// it is written as four aligned float4; x-halo cells at [4-H, 4), [20, 20+H).
constexpr int kPitch = 24;

// Element storage of a fine slab: FP32, or Density (u16, channel.rs:118-142)
// loaded as u * (1/65535) (Density::to_f32, the halo's dequant, halo.rs:95)
// and stored as round(v * 65535), half away from zero (StoreBack<Density>,
// math/stencil.rs:162-174), saturated to the u16 range.
__device__ __forceinline__ float deq(float v) { return v; }
__device__ __forceinline__ float deq(uint16_t v) { return (float)v * (1.0f / 65535.0f); }
__device__ __forceinline__ float4 load4(const float *p) { return *reinterpret_cast<const float4 *>(p); }
__device__ __forceinline__ float4 load4(const uint16_t *p) {
    const uint2 u = *reinterpret_cast<const uint2 *>(p);
    return make_float4(deq((uint16_t)(u.x & 0xffffu)), deq((uint16_t)(u.x >> 16)), deq((uint16_t)(u.y & 0xffffu)),
                       deq((uint16_t)(u.y >> 16)));
}
__device__ __forceinline__ void store_back(float *p, float v) { *p = v; }
__device__ __forceinline__ void store_back(uint16_t *p, float v) {
    const float r = fminf(fmaxf(roundf(v * 65535.0f), 0.0f), 65535.0f);
    *p = (uint16_t)(int)r;
}

// ---------------------------------------------------------------------------
// Halo topology.  Entry t of a block's row (t = dx + n*dy + n*n*dz over the
// (2H+1)^3 neighbourhood) is the neighbour's fine slot (>= 0; 0 = empty dummy,
// also outside the domain) or -(coarse slot) - 1 for a coarser neighbour.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int neighbour_code(const GridView &g, const GridView &c, const uint8_t *levels,
                                              int64_t bid, int bx, int by, int bz) {
    if (bx < 0 || bx >= g.nx || by < 0 || by >= g.ny || bz < 0 || bz >= g.nz) return 0;
    const int64_t b = (int64_t)bx + (int64_t)by * g.nx + (int64_t)bz * g.nx * g.ny;
    if (levels && levels[b] > levels[bid]) return -c.b2s[b] - 1;
    return g.b2s[b];
}

template <int HALO>
__global__ void topology_kernel(GridView g, GridView c, const uint8_t *levels, const int64_t *bids, int64_t n,
                                int32_t *topo) {
    constexpr int NN = 2 * HALO + 1, NB = NN * NN * NN;
    const int64_t i = blockIdx.x;
    const int t = threadIdx.x;
    if (i >= n || t >= NB) return;
    const int64_t bid = bids[i];
    const int bx0 = (int)(bid % g.nx), by0 = (int)((bid / g.nx) % g.ny), bz0 = (int)(bid / ((int64_t)g.nx * g.ny));
    const int dx = t % NN, dy = (t / NN) % NN, dz = t / (NN * NN);
    topo[i * NB + t] = neighbour_code(g, c, levels, bid, bx0 + dx - HALO, by0 + dy - HALO, bz0 + dz - HALO);
}

// Same resolution inline (standalone gather, which has no plan).
template <int HALO, bool MULTI>
__device__ __forceinline__ void resolve_neighbours(const GridView &g, const GridView &c, const uint8_t *levels,
                                                   int bx0, int by0, int bz0, int64_t bid, int *s_nb,
                                                   int *s_nbc) {
    constexpr int n = 2 * HALO + 1;
    const int t = threadIdx.x;
    if (t < n * n * n) {
        const int code = neighbour_code(g, c, MULTI ? levels : nullptr, bid, bx0 + t % n - HALO,
                                        by0 + (t / n) % n - HALO, bz0 + t / (n * n) - HALO);
        s_nb[t] = code >= 0 ? code : 0;
        if (MULTI) s_nbc[t] = code >= 0 ? -1 : -code - 1;
    }
}

// ---------------------------------------------------------------------------
// Tile fills (halo.rs:31-170, fill_multires halo.rs:187-358).  A tile cell
// (hx, hy, hz) holds grid cell (x0 + hx - H, ...), each axis resolved by the
// boundary condition; a coarse neighbour is sampled at (x>>1, y>>1, z>>1) in
// its 8^3 block.
// ---------------------------------------------------------------------------
// (The coarse level is FP32 only: a coarse pointer is returned as TS = float.)
template <int HALO, bool MULTI, typename TS>
__device__ __forceinline__ const TS *cell_ptr(const TS *fs, const GridView &c, const int *s_nb, const int *s_nbc,
                                              int bx0, int by0, int bz0, int rx, int ry, int rz) {
    constexpr int n = 2 * HALO + 1;
    const int nb = ((rx >> 4) - bx0 + HALO) + n * ((ry >> 4) - by0 + HALO) + n * n * ((rz >> 4) - bz0 + HALO);
    if (MULTI && s_nbc[nb] >= 0)
        return reinterpret_cast<const TS *>(c.slab + (int64_t)s_nbc[nb] * 512 +
                                            (((rx >> 1) & 7) | (((ry >> 1) & 7) << 3) | (((rz >> 1) & 7) << 6)));
    return fs + (int64_t)s_nb[nb] * 4096 + ((rx & 15) | ((ry & 15) << 4) | ((rz & 15) << 8));
}

// Gather tile, dense HSX^3 layout.  Every lane resolves all its cells first
// and then issues its global loads back to back.  `full == false` writes 0
// where a faces-only fill leaves a fresh HaloBlock untouched (halo.rs:154).
template <int HALO, bool MULTI>
__device__ __forceinline__ void fill_tile_dense(float *tile, const GridView &g, const GridView &c, const int *s_nb,
                                                const int *s_nbc, int bx0, int by0, int bz0, int bc, float empty_f32,
                                                bool full) {
    constexpr int B = 16, HSX = B + 2 * HALO, NT = HSX * HSX * HSX;
    constexpr int ITEMS = (NT + kGatherThreads - 1) / kGatherThreads;
    const int x0 = bx0 * B, y0 = by0 * B, z0 = bz0 * B;
    const float *src[ITEMS];
    uint32_t hole = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const int i = threadIdx.x + k * kGatherThreads;
        src[k] = nullptr;
        if (i >= NT) continue;
        const int hx = i % HSX, hy = (i / HSX) % HSX, hz = i / (HSX * HSX);
        const bool xh = hx < HALO || hx >= HALO + B, yh = hy < HALO || hy >= HALO + B, zh = hz < HALO || hz >= HALO + B;
        if (!full && xh && (yh || zh)) {
            hole |= 1u << k;
            continue;
        }
        int rx, ry, rz;
        if (resolve_axis(x0 + hx - HALO, g.sx - 1, bc, rx) && resolve_axis(y0 + hy - HALO, g.sy - 1, bc, ry) &&
            resolve_axis(z0 + hz - HALO, g.sz - 1, bc, rz))
            src[k] = cell_ptr<HALO, MULTI, float>(g.slab, c, s_nb, s_nbc, bx0, by0, bz0, rx, ry, rz);
        // else Dirichlet outside the domain: the empty sentinel
    }
    float v[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) v[k] = src[k] ? *src[k] : ((hole >> k) & 1u ? 0.0f : empty_f32);
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const int i = threadIdx.x + k * kGatherThreads;
        if (i < NT) tile[i] = v[k];
    }
}

template <int HALO>
__device__ __forceinline__ int tile_at(int hx, int hy, int hz) {
    constexpr int HSX = 16 + 2 * HALO;
    return (hz * HSX + hy) * kPitch + (4 - HALO) + hx;
}

// Pitched per-cell fill: the path for a block whose rows cross the domain's
// x end (partial last block, halo.rs:130-144).
template <int HALO, bool MULTI, typename TS>
__device__ void fill_tile_pitched_generic(float *tile, const TS *fs, const GridView &g, const GridView &c,
                                          const int *s_nb, const int *s_nbc, int bx0, int by0, int bz0,
                                          float empty_f32) {
    constexpr int B = 16, HSX = B + 2 * HALO, NT = HSX * HSX * HSX;
    const int x0 = bx0 * B, y0 = by0 * B, z0 = bz0 * B;
    for (int i = threadIdx.x; i < NT; i += blockDim.x) {
        const int hx = i % HSX, hy = (i / HSX) % HSX, hz = i / (HSX * HSX);
        int rx, ry, rz;
        float v = empty_f32;
        if (resolve_axis(x0 + hx - HALO, g.sx - 1, MSBG_BC_NEUMANN, rx) &&
            resolve_axis(y0 + hy - HALO, g.sy - 1, MSBG_BC_NEUMANN, ry) &&
            resolve_axis(z0 + hz - HALO, g.sz - 1, MSBG_BC_NEUMANN, rz))
            v = deq(*cell_ptr<HALO, MULTI, TS>(fs, c, s_nb, s_nbc, bx0, by0, bz0, rx, ry, rz));
        tile[tile_at<HALO>(hx, hy, hz)] = v;
    }
}

// Row-wise pitched fill for blocks whose rows lie inside the domain in x.  A
// tile row (hy, hz) is resolved once: its 16 interior cells are one block row
// (four float4 loads, or two float2 loads duplicated for a coarse
// neighbour), its 2H x-halo cells are loaded one by one.  `skip_edges` leaves
// out the rows and x-halo cells a 7-point stencil never reads.
template <int HALO, bool MULTI, int NTHR, typename TS>
__device__ __forceinline__ void fill_tile_rows(float *tile, const TS *fs, const GridView &g, const GridView &c,
                                               const int *s_nb, const int *s_nbc, int bx0, int by0, int bz0,
                                               float empty_f32, bool skip_edges) {
    constexpr int B = 16, L = 4, HSX = B + 2 * HALO, n = 2 * HALO + 1;
    constexpr int NR = HSX * HSX, NQ = 4 * NR, NH = 2 * HALO * NR;
    constexpr int QI = (NQ + NTHR - 1) / NTHR, HI = (NH + NTHR - 1) / NTHR;
    const int x0 = bx0 * B, y0 = by0 * B, z0 = bz0 * B;
    const int t = threadIdx.x;
    const TS *qp[QI];
    uint32_t qcoarse = 0, qwrite = 0;
#pragma unroll
    for (int k = 0; k < QI; ++k) {  // interior quads: item = row * 4 + q
        const int it = t + k * NTHR;
        qp[k] = nullptr;
        if (it >= NQ) continue;
        const int r = it >> 2, q = it & 3;
        const int hy = r % HSX, hz = r / HSX;
        const bool yh = hy < HALO || hy >= HALO + B, zh = hz < HALO || hz >= HALO + B;
        if (skip_edges && yh && zh) continue;
        qwrite |= 1u << k;
        int ry, rz;
        if (!resolve_axis(y0 + hy - HALO, g.sy - 1, MSBG_BC_NEUMANN, ry) ||
            !resolve_axis(z0 + hz - HALO, g.sz - 1, MSBG_BC_NEUMANN, rz))
            continue;
        const int nb = HALO + n * ((ry >> L) - by0 + HALO) + n * n * ((rz >> L) - bz0 + HALO);
        if (MULTI && s_nbc[nb] >= 0) {
            qcoarse |= 1u << k;
            qp[k] = reinterpret_cast<const TS *>(c.slab + (int64_t)s_nbc[nb] * 512 + (((ry >> 1) & 7) << 3) +
                                                 (((rz >> 1) & 7) << 6) + 2 * q);
        } else {
            qp[k] = fs + (int64_t)s_nb[nb] * 4096 + ((ry & 15) << 4) + ((rz & 15) << 8) + 4 * q;
        }
    }
    const TS *hp[HI];
    uint32_t hwrite = 0;
#pragma unroll
    for (int k = 0; k < HI; ++k) {  // x-halo cells: item = row * 2H + j; j < H left
        const int it = t + k * NTHR;
        hp[k] = nullptr;
        if (it >= NH) continue;
        const int r = it / (2 * HALO), j = it % (2 * HALO);
        const int hy = r % HSX, hz = r / HSX;
        const bool yh = hy < HALO || hy >= HALO + B, zh = hz < HALO || hz >= HALO + B;
        if (skip_edges && (yh || zh)) continue;
        hwrite |= 1u << k;
        const int gx = j < HALO ? x0 - HALO + j : x0 + B + (j - HALO);
        int rx, ry, rz;
        if (resolve_axis(gx, g.sx - 1, MSBG_BC_NEUMANN, rx) &&
            resolve_axis(y0 + hy - HALO, g.sy - 1, MSBG_BC_NEUMANN, ry) &&
            resolve_axis(z0 + hz - HALO, g.sz - 1, MSBG_BC_NEUMANN, rz))
            hp[k] = cell_ptr<HALO, MULTI, TS>(fs, c, s_nb, s_nbc, bx0, by0, bz0, rx, ry, rz);
    }
    float4 qv[QI];
#pragma unroll
    for (int k = 0; k < QI; ++k) {
        if (!qp[k]) {
            qv[k] = make_float4(empty_f32, empty_f32, empty_f32, empty_f32);
        } else if (MULTI && ((qcoarse >> k) & 1u)) {
            const float2 c2 = *reinterpret_cast<const float2 *>(qp[k]);
            qv[k] = make_float4(c2.x, c2.x, c2.y, c2.y);
        } else {
            qv[k] = load4(qp[k]);
        }
    }
    float hv[HI];
#pragma unroll
    for (int k = 0; k < HI; ++k) hv[k] = hp[k] ? deq(*hp[k]) : empty_f32;
#pragma unroll
    for (int k = 0; k < QI; ++k)
        if ((qwrite >> k) & 1u) {
            const int it = t + k * NTHR, r = it >> 2, q = it & 3;
            *reinterpret_cast<float4 *>(tile + r * kPitch + 4 + 4 * q) = qv[k];
        }
#pragma unroll
    for (int k = 0; k < HI; ++k)
        if ((hwrite >> k) & 1u) {
            const int it = t + k * NTHR, r = it / (2 * HALO), j = it % (2 * HALO);
            tile[r * kPitch + (j < HALO ? 4 - HALO + j : 20 + (j - HALO))] = hv[k];
        }
}

__device__ __forceinline__ float clamp_f(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

struct Cross {
    float c, xm, xp, ym, yp;
};

// One colour pass: workgroup k updates block bids[k] in place, its halo
// topology in topo[k * (2H+1)^3 ...].  512 lanes = 16 x 16 (x, y) x 2
// z-halves; each lane walks 8 cells up its z column keeping the planes it
// still needs in registers (5 / 9 / 13 LDS reads per cell for the 7 / 19 /
// 25-tap stencils once the window is primed).
template <int STENCIL, bool MULTI, typename TS>
__global__ __launch_bounds__(kSweepThreads) void sweep_kernel(GridView g, GridView c, TS *__restrict__ fslab,
                                                              const float *__restrict__ cslab, const int64_t *bids,
                                                              const int32_t *topo, float dt, int clamp01,
                                                              float empty_f32) {
    static_assert(!MULTI || sizeof(TS) == sizeof(float), "cross-level sweeps are FP32 only");
    // The slabs come in as plain pointer arguments so every access compiles
    // to a global (not flat) memory instruction.
    c.slab = const_cast<float *>(cslab);
    constexpr int HALO = STENCIL == MSBG_BILAPLACIAN ? 2 : 1;
    constexpr int HSX = 16 + 2 * HALO, NN = 2 * HALO + 1, NB = NN * NN * NN, CENTER = (NB - 1) / 2;
    constexpr int CPT = 16 * 256 / kSweepThreads;  // cells per lane along z
    __shared__ __attribute__((aligned(16))) float tile[HSX * HSX * kPitch];
    __shared__ int s_nb[NB];
    __shared__ int s_nbc[MULTI ? NB : 1];
    const int wg = xcd_remap(blockIdx.x, gridDim.x);
    const int32_t *row = topo + (int64_t)wg * NB;
    const int slot = row[CENTER];
    if (slot < 2) return;  // dummy or absent: solver.rs:201-204 (uniform per workgroup)
    const int64_t bid = bids[wg];
    if (threadIdx.x < NB) {
        const int code = row[threadIdx.x];
        s_nb[threadIdx.x] = code >= 0 ? code : 0;
        if (MULTI) s_nbc[threadIdx.x] = code >= 0 ? -1 : -code - 1;
    }
    const int bx0 = (int)(bid % g.nx), by0 = (int)((bid / g.nx) % g.ny), bz0 = (int)(bid / ((int64_t)g.nx * g.ny));
    __syncthreads();
    if (bx0 * 16 + 16 <= g.sx)
        fill_tile_rows<HALO, MULTI, kSweepThreads, TS>(tile, fslab, g, c, s_nb, s_nbc, bx0, by0, bz0, empty_f32,
                                                   STENCIL == MSBG_LAPLACIAN);
    else
        fill_tile_pitched_generic<HALO, MULTI, TS>(tile, fslab, g, c, s_nb, s_nbc, bx0, by0, bz0, empty_f32);
    __syncthreads();

    const int x = threadIdx.x & 15, y = (threadIdx.x >> 4) & 15, zs = (threadIdx.x >> 8) * CPT;
    TS *out = fslab + (int64_t)slot * 4096 + y * 16 + x;
    const bool cl = clamp01 != 0;
    // T(dx, dy, z): tile value at block cell (x+dx, y+dy, z).
    auto T = [&](int dx, int dy, int z) -> float {
        return tile[tile_at<HALO>(x + HALO + dx, y + HALO + dy, z + HALO)];
    };
    auto cross = [&](int z) -> Cross { return Cross{T(0, 0, z), T(-1, 0, z), T(1, 0, z), T(0, -1, z), T(0, 1, z)}; };

    if (STENCIL == MSBG_LAPLACIAN) {
        // kernel_laplacian -- math/laplacian.rs:44-61
        float cm = T(0, 0, zs - 1), c0 = T(0, 0, zs);
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            const int z = zs + k;
            const float cp = T(0, 0, z + 1);
            const float cc = c0, lx = T(-1, 0, z), rx = T(1, 0, z), dyu = T(0, -1, z), dyd = T(0, 1, z), dzb = cm,
                        dzf = cp;
            const float nb = lx + rx + dyu + dyd + dzb + dzf;
            const float lap = __builtin_fmaf(cc, -6.0f, nb);
            const float fx = (rx - lx) * 0.5f, fy = (dyd - dyu) * 0.5f, fz = (dzf - dzb) * 0.5f;
            const float gm = __builtin_sqrtf(__builtin_fmaf(fx, fx, __builtin_fmaf(fy, fy, fz * fz)));
            const float v = __builtin_fmaf(dt * lap, gm, cc);
            store_back(out + z * 256, cl ? clamp_f(v, 0.0f, 1.0f) : v);
            cm = c0;
            c0 = cp;
        }
    } else if (STENCIL == MSBG_MEAN_CURVATURE) {
        // kernel_meancurv -- math/meancurv.rs:49-100
        Cross pm = cross(zs - 1), p0 = cross(zs);
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            const int z = zs + k;
            const Cross pp = cross(z + 1);
            const float f0 = p0.c, f1 = p0.xm, f2 = p0.xp, f3 = p0.ym, f4 = p0.yp, f5 = pm.c, f6 = pp.c;
            const float fx = (f2 - f1) * 0.5f, fy = (f4 - f3) * 0.5f, fz = (f6 - f5) * 0.5f;
            const float fxx = f1 + f2 - 2.0f * f0, fyy = f3 + f4 - 2.0f * f0, fzz = f5 + f6 - 2.0f * f0;
            const float d1 = T(1, 1, z), d2 = T(1, -1, z), d3 = T(-1, 1, z), d4 = T(-1, -1, z);
            const float d5 = pp.xp, d6 = pm.xp, d7 = pp.xm, d8 = pm.xm;
            const float d9 = pp.yp, d10 = pm.yp, d11 = pp.ym, d12 = pm.ym;
            const float fyx = (d1 - d2 - d3 + d4) * 0.25f;
            const float fzx = (d5 - d6 - d7 + d8) * 0.25f;
            const float fzy = (d9 - d10 - d11 + d12) * 0.25f;
            const float crs = fx * (fy * fyx) + fx * (fz * fzx) + fy * (fz * fzy);
            const float hnum = __builtin_fmaf(fy * fy + fz * fz, fxx,
                                              __builtin_fmaf(fx * fx + fz * fz, fyy,
                                                             __builtin_fmaf(fx * fx + fy * fy, fzz, -2.0f * crs)));
            const float grad = __builtin_fmaf(fx, fx, __builtin_fmaf(fy, fy, fz * fz));
            const float hc = grad > 1e-7f ? hnum / grad : 0.0f;
            const float d = clamp_f(dt * hc, -0.1f, 0.1f);
            const float v = f0 + d;
            store_back(out + z * 256, cl ? clamp_f(v, 0.0f, 1.0f) : v);
            pm = p0;
            p0 = pp;
        }
    } else {
        // kernel_bilaplacian -- math/bilaplacian.rs:47-88
        float cm2 = T(0, 0, zs - 2);
        Cross pm = cross(zs - 1), p0 = cross(zs), pp = cross(zs + 1);
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            const int z = zs + k;
            const Cross pp2 = cross(z + 2);
            const float f0 = p0.c, f1 = p0.xm, f2 = p0.xp, f3 = p0.ym, f4 = p0.yp, f5 = pm.c, f6 = pp.c;
            const float s1 = f1 + f2 + f3 + f4 + f5 + f6;
            const float s2 = T(-2, 0, z) + T(2, 0, z) + T(0, -2, z) + T(0, 2, z) + cm2 + pp2.c;
            const float sd = T(1, 1, z) + T(1, -1, z) + T(-1, 1, z) + T(-1, -1, z) + pp.xp + pm.xp + pp.xm + pm.xm +
                             pp.yp + pm.yp + pp.ym + pm.ym;
            const float hh = 42.0f * f0 - 12.0f * s1 + s2 + 2.0f * sd;
            const float fx = (f2 - f1) * 0.5f, fy = (f4 - f3) * 0.5f, fz = (f6 - f5) * 0.5f;
            const float gm = __builtin_sqrtf(__builtin_fmaf(fx, fx, __builtin_fmaf(fy, fy, fz * fz)));
            const float upd = (dt * hh) * gm;
            const float d = clamp_f(-upd, -0.1f, 0.1f);
            const float v = f0 + d;
            store_back(out + z * 256, cl ? clamp_f(v, 0.0f, 1.0f) : v);
            cm2 = pm.c;
            pm = p0;
            p0 = pp;
            pp = pp2;
        }
    }
}

// Standalone halo gather: workgroup k writes the dense tile of bids[k].
template <int HALO, bool MULTI>
__global__ __launch_bounds__(kGatherThreads) void gather_kernel(GridView g, GridView c, const uint8_t *levels,
                                                                const int64_t *bids, int bc, int full,
                                                                float empty_f32, float *tiles) {
    constexpr int HSX = 16 + 2 * HALO, NT = HSX * HSX * HSX;
    __shared__ float tile[NT];
    __shared__ int s_nb[125];
    __shared__ int s_nbc[MULTI ? 125 : 1];
    const int64_t bid = bids[blockIdx.x];
    const int bx0 = (int)(bid % g.nx), by0 = (int)((bid / g.nx) % g.ny), bz0 = (int)(bid / ((int64_t)g.nx * g.ny));
    resolve_neighbours<HALO, MULTI>(g, c, levels, bx0, by0, bz0, bid, s_nb, s_nbc);
    __syncthreads();
    fill_tile_dense<HALO, MULTI>(tile, g, c, s_nb, s_nbc, bx0, by0, bz0, bc, empty_f32, full != 0);
    __syncthreads();
    float *dst = tiles + (int64_t)blockIdx.x * NT;
    for (int i = threadIdx.x; i < NT; i += kGatherThreads) dst[i] = tile[i];
}

}  // namespace msbg

using namespace msbg;

struct msbg_sweep_plan {
    msbg_grid *g = nullptr;
    const msbg_grid *coarse = nullptr;
    const uint8_t *d_levels = nullptr;
    int64_t *d_bids = nullptr;
    int32_t *d_topo[3] = {nullptr, nullptr, nullptr};  // by halo width 1, 2
    int64_t counts[8] = {0};
    int64_t n = 0;
    uint16_t *slab16 = nullptr;  // Density (u16) storage over g's layout, or null (FP32)
    // Cached graph of the last replayed sweep (run_graph).
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t exec = nullptr;
    int g_stencil = -1, g_iters = -1, g_clamp = -1;
    float g_dt = 0.0f;
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

int check_levels(const msbg_grid *g, const msbg_grid *coarse, const uint8_t *d_levels) {
    MSBG_CHECK(coarse->L == g->L - 1 && coarse->nx == g->nx && coarse->ny == g->ny && coarse->nz == g->nz, MSBG_EINVAL,
               "multires: the coarse level must have half the block edge over the same block grid");
    MSBG_CHECK(d_levels, MSBG_EINVAL, "multires: null level map");
    return MSBG_OK;
}

// Topology rows of `n` device bids for halo width `halo`.
int build_topology(const msbg_grid *g, const msbg_grid *coarse, const uint8_t *d_levels, const int64_t *d_bids,
                   int64_t n, int halo, int32_t *d_topo) {
    if (n == 0) return MSBG_OK;
    GridView gv = g->view();
    GridView cv = coarse ? coarse->view() : gv;
    const uint8_t *lv = coarse ? d_levels : nullptr;
    if (halo == 1)
        hipLaunchKernelGGL(topology_kernel<1>, dim3((unsigned)n), dim3(32), 0, stream(), gv, cv, lv, d_bids, n, d_topo);
    else
        hipLaunchKernelGGL(topology_kernel<2>, dim3((unsigned)n), dim3(128), 0, stream(), gv, cv, lv, d_bids, n,
                           d_topo);
    MSBG_HIP(hipGetLastError());
    return MSBG_OK;
}

int plan_new(msbg_grid *g, const msbg_grid *coarse, const uint8_t *d_levels, const int64_t *active, int64_t n,
             msbg_sweep_plan **out) {
    MSBG_CHECK(out, MSBG_EINVAL, "null out");
    *out = nullptr;
    MSBG_CHECK(g && g->L == 4, MSBG_EINVAL, "sweeps run on 16^3-block grids");
    if (coarse) {
        int rc = check_levels(g, coarse, d_levels);
        if (rc) return rc;
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
        if (e == hipSuccess) e = hipMalloc(&p->d_topo[1], sizeof(int32_t) * 27 * n);
        if (e == hipSuccess) e = hipMalloc(&p->d_topo[2], sizeof(int32_t) * 125 * n);
        if (e == hipSuccess)
            e = hipMemcpyAsync(p->d_bids, order.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, stream());
        if (e != hipSuccess) {
            msbg_sweep_plan_free(p);
            return hip_fail(e, "sweep plan upload");
        }
        rc = build_topology(g, coarse, d_levels, p->d_bids, n, 1, p->d_topo[1]);
        if (rc == MSBG_OK) rc = build_topology(g, coarse, d_levels, p->d_bids, n, 2, p->d_topo[2]);
        if (rc == MSBG_OK) {
            e = hipStreamSynchronize(stream());
            if (e != hipSuccess) rc = hip_fail(e, "sweep plan topology");
        }
        if (rc) {
            msbg_sweep_plan_free(p);
            return rc;
        }
    }
    *out = p;
    return MSBG_OK;
}

template <bool MULTI, typename TS>
void launch_colour(int stencil, dim3 grid, hipStream_t s, GridView g, GridView c, TS *fslab, const int64_t *bids,
                   const int32_t *topo, float dt, int clamp01, float empty_f32) {
    const dim3 block(kSweepThreads);
    switch (stencil) {
        case MSBG_LAPLACIAN:
            hipLaunchKernelGGL((sweep_kernel<MSBG_LAPLACIAN, MULTI, TS>), grid, block, 0, s, g, c, fslab, c.slab, bids,
                               topo, dt, clamp01, empty_f32);
            break;
        case MSBG_MEAN_CURVATURE:
            hipLaunchKernelGGL((sweep_kernel<MSBG_MEAN_CURVATURE, MULTI, TS>), grid, block, 0, s, g, c, fslab, c.slab,
                               bids, topo, dt, clamp01, empty_f32);
            break;
        default:
            hipLaunchKernelGGL((sweep_kernel<MSBG_BILAPLACIAN, MULTI, TS>), grid, block, 0, s, g, c, fslab, c.slab,
                               bids, topo, dt, clamp01, empty_f32);
    }
}

// Launches the 8 colour passes of one sweep on `s`.  When `ev` is non-null,
// an event pair brackets every launch.
template <bool MULTI>
int launch_sweep(msbg_sweep_plan *p, hipStream_t s, int stencil, float dt, int clamp01,
                 std::vector<hipEvent_t> *ev = nullptr) {
    GridView g = p->g->view();
    GridView c = MULTI ? p->coarse->view() : g;
    const float empty_f32 = MULTI ? 0.0f : (p->slab16 ? 0.0f : p->g->empty_value);
    const int halo = stencil == MSBG_BILAPLACIAN ? 2 : 1;
    const int nb = halo == 2 ? 125 : 27;
    int64_t off = 0;
    for (int col = 0; col < 8; ++col) {
        const int64_t cnt = p->counts[col];
        if (cnt > 0) {
            const dim3 grid((unsigned)cnt);
            const int64_t *bids = p->d_bids + off;
            const int32_t *topo = p->d_topo[halo] + off * nb;
            if (ev) {
                hipEvent_t a, b;
                MSBG_HIP(hipEventCreate(&a));
                MSBG_HIP(hipEventCreate(&b));
                ev->push_back(a);
                ev->push_back(b);
                MSBG_HIP(hipEventRecord(a, s));
            }
            if (!MULTI && p->slab16)
                launch_colour<false, uint16_t>(stencil, grid, s, g, c, p->slab16, bids, topo, dt, clamp01, empty_f32);
            else
                launch_colour<MULTI, float>(stencil, grid, s, g, c, g.slab, bids, topo, dt, clamp01, empty_f32);
            MSBG_HIP(hipGetLastError());
            if (ev) MSBG_HIP(hipEventRecord(ev->back(), s));
        }
        off += cnt;
    }
    return MSBG_OK;
}

int launch_iterations(msbg_sweep_plan *p, hipStream_t s, int stencil, float dt, int iterations, int clamp01,
                      std::vector<hipEvent_t> *ev = nullptr) {
    for (int it = 0; it < iterations; ++it) {
        int rc = p->coarse ? launch_sweep<true>(p, s, stencil, dt, clamp01, ev)
                           : launch_sweep<false>(p, s, stencil, dt, clamp01, ev);
        if (rc) return rc;
    }
    return MSBG_OK;
}

void drop_graph(msbg_sweep_plan *p) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    p->exec = nullptr;
    p->g_stencil = -1;
}

int gather(const msbg_grid *g, const msbg_grid *coarse, const uint8_t *d_levels, const int64_t *bids, int64_t n,
           int halo, int full, int bc, float *d_tiles) {
    MSBG_CHECK(g && g->L == 4, MSBG_EINVAL, "halo gather runs on 16^3-block grids");
    MSBG_CHECK(halo == 1 || halo == 2, MSBG_EINVAL, "halo must be 1 or 2 (halo.rs:62)");
    MSBG_CHECK(bc >= 0 && bc <= 2, MSBG_EINVAL, "unknown boundary condition");
    MSBG_CHECK(n >= 0 && (n == 0 || (bids && d_tiles)), MSBG_EINVAL, "halo gather: null argument");
    if (coarse) {
        int rc = check_levels(g, coarse, d_levels);
        if (rc) return rc;
    }
    for (int64_t i = 0; i < n; ++i)
        MSBG_CHECK(bids[i] >= 0 && bids[i] < g->n_blocks, MSBG_EINVAL, "halo gather: bid out of range");
    if (n == 0) return MSBG_OK;
    int rc = MSBG_OK;
    // the block list goes to a cached device buffer (no allocation per call)
    int64_t *d_bids = (int64_t *)workspace(21, sizeof(int64_t) * n, &rc);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(d_bids, bids, sizeof(int64_t) * n, hipMemcpyHostToDevice, stream());
    if (e == hipSuccess) {
        GridView gv = g->view();
        GridView cv = coarse ? coarse->view() : gv;
        const float empty_f32 = coarse ? 0.0f : g->empty_value;
        dim3 grid((unsigned)n), block(kGatherThreads);
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

int msbg_sweep_plan_new_density(msbg_grid *layout, uint16_t *d_slab16, const int64_t *active, int64_t n,
                                msbg_sweep_plan **out) {
    MSBG_CHECK(d_slab16, MSBG_EINVAL, "density plan: null u16 slab");
    int rc = plan_new(layout, nullptr, nullptr, active, n, out);
    if (rc) return rc;
    (*out)->slab16 = d_slab16;
    return MSBG_OK;
}

int msbg_sweep_density(msbg_grid *layout, uint16_t *d_slab16, const int64_t *active, int64_t n, int stencil, float dt,
                       int iterations, int clamp01) {
    int rc = check_stencil(stencil);
    if (rc) return rc;
    msbg_sweep_plan *p = nullptr;
    rc = msbg_sweep_plan_new_density(layout, d_slab16, active, n, &p);
    if (rc) return rc;
    rc = msbg_sweep_plan_run_async(p, stencil, dt, iterations, clamp01);
    if (rc == MSBG_OK) {
        hipError_t e = hipStreamSynchronize(stream());
        if (e != hipSuccess) rc = hip_fail(e, "msbg_sweep_density");
    }
    msbg_sweep_plan_free(p);
    return rc;
}

int msbg_sweep_plan_new_multires(msbg_grid *fine, const msbg_grid *coarse, const uint8_t *dev_levels,
                                 const int64_t *active, int64_t n, msbg_sweep_plan **out) {
    MSBG_CHECK(coarse, MSBG_EINVAL, "multires plan: null coarse level");
    return plan_new(fine, coarse, dev_levels, active, n, out);
}

int msbg_sweep_plan_free(msbg_sweep_plan *p) {
    if (!p) return MSBG_OK;
    drop_graph(p);
    if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
    if (p->d_bids) (void)hipFree(p->d_bids);
    for (int32_t *t : p->d_topo)
        if (t) (void)hipFree(t);
    delete p;
    return MSBG_OK;
}

int msbg_sweep_plan_run_async(msbg_sweep_plan *p, int stencil, float dt, int iterations, int clamp01) {
    MSBG_CHECK(p, MSBG_EINVAL, "null plan");
    int rc = check_stencil(stencil);
    if (rc) return rc;
    MSBG_CHECK(iterations >= 0, MSBG_EINVAL, "negative iterations");
    return launch_iterations(p, stream(), stencil, dt, iterations, clamp01);
}

int msbg_sweep_plan_run_graph_async(msbg_sweep_plan *p, int stencil, float dt, int iterations, int clamp01) {
    MSBG_CHECK(p, MSBG_EINVAL, "null plan");
    int rc = check_stencil(stencil);
    if (rc) return rc;
    MSBG_CHECK(iterations >= 0, MSBG_EINVAL, "negative iterations");
    if (iterations == 0 || p->n == 0) return MSBG_OK;
    if (!(p->exec && p->g_stencil == stencil && p->g_iters == iterations && p->g_clamp == clamp01 && p->g_dt == dt)) {
        drop_graph(p);
        if (!p->cap_stream) MSBG_HIP(hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking));
        MSBG_HIP(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeRelaxed));
        rc = launch_iterations(p, p->cap_stream, stencil, dt, iterations, clamp01);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(p->cap_stream, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        MSBG_HIP(e);
        e = hipGraphInstantiate(&p->exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        MSBG_HIP(e);
        p->g_stencil = stencil;
        p->g_iters = iterations;
        p->g_clamp = clamp01;
        p->g_dt = dt;
    }
    MSBG_HIP(hipGraphLaunch(p->exec, stream()));
    return MSBG_OK;
}

int msbg_sweep_plan_run_timed(msbg_sweep_plan *p, int stencil, float dt, int iterations, int clamp01,
                              float *kernel_ms, int64_t *launches) {
    MSBG_CHECK(p && kernel_ms && launches, MSBG_EINVAL, "null argument");
    int rc = check_stencil(stencil);
    if (rc) return rc;
    MSBG_CHECK(iterations >= 0, MSBG_EINVAL, "negative iterations");
    std::vector<hipEvent_t> ev;
    rc = launch_iterations(p, stream(), stencil, dt, iterations, clamp01, &ev);
    float total = 0.0f;
    hipError_t e = hipStreamSynchronize(stream());
    for (size_t k = 0; k + 1 < ev.size() && e == hipSuccess; k += 2) {
        float ms = 0.0f;
        e = hipEventElapsedTime(&ms, ev[k], ev[k + 1]);
        total += ms;
    }
    for (hipEvent_t x : ev) (void)hipEventDestroy(x);
    if (rc) return rc;
    MSBG_HIP(e);
    *kernel_ms = total;
    *launches = (int64_t)(ev.size() / 2);
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

int msbg_halo_topology(const msbg_grid *g, const msbg_grid *coarse, const uint8_t *dev_levels, const int64_t *bids,
                       int64_t n, int halo, int32_t *host_out) {
    MSBG_CHECK(g, MSBG_EINVAL, "halo topology: null grid");
    MSBG_CHECK(halo == 1 || halo == 2, MSBG_EINVAL, "halo must be 1 or 2 (halo.rs:62)");
    MSBG_CHECK(n >= 0 && (n == 0 || (bids && host_out)), MSBG_EINVAL, "halo topology: null argument");
    if (coarse) {
        int rc = check_levels(g, coarse, dev_levels);
        if (rc) return rc;
    }
    for (int64_t i = 0; i < n; ++i)
        MSBG_CHECK(bids[i] >= 0 && bids[i] < g->n_blocks, MSBG_EINVAL, "halo topology: bid out of range");
    if (n == 0) return MSBG_OK;
    const int nb = halo == 2 ? 125 : 27;
    int64_t *d_bids = nullptr;
    int32_t *d_topo = nullptr;
    MSBG_HIP(hipMalloc(&d_bids, sizeof(int64_t) * n));
    hipError_t e = hipMalloc(&d_topo, sizeof(int32_t) * nb * n);
    if (e == hipSuccess) e = hipMemcpyAsync(d_bids, bids, sizeof(int64_t) * n, hipMemcpyHostToDevice, stream());
    int rc = MSBG_OK;
    if (e == hipSuccess) rc = build_topology(g, coarse, dev_levels, d_bids, n, halo, d_topo);
    if (e == hipSuccess && rc == MSBG_OK)
        e = hipMemcpyAsync(host_out, d_topo, sizeof(int32_t) * nb * n, hipMemcpyDeviceToHost, stream());
    if (e == hipSuccess) e = hipStreamSynchronize(stream());
    (void)hipFree(d_bids);
    if (d_topo) (void)hipFree(d_topo);
    if (rc) return rc;
    MSBG_HIP(e);
    return MSBG_OK;
}

}  // extern "C"

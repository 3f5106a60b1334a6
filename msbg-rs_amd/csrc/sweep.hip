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

#include <type_traits>

#include "density8.hpp"
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
// Slab reads through the global address space: the slabs are device
// allocations, and a global load (unlike a flat one) leaves the LDS counter
// alone, so the wait for one load does not also wait for LDS traffic.
template <typename T>
__device__ __forceinline__ T ld_global(const void *p) {
    return *(const __attribute__((address_space(1))) T *)p;
}
typedef float gf4 __attribute__((ext_vector_type(4)));  // native vectors: HIP's float4 copies through a generic reference
typedef uint32_t gu2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float4 load4(const float *p) {
    const gf4 v = ld_global<gf4>(p);
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float4 load4(const uint16_t *p) {
    const gu2 u = ld_global<gu2>(p);
    return make_float4(deq((uint16_t)(u.x & 0xffffu)), deq((uint16_t)(u.x >> 16)), deq((uint16_t)(u.y & 0xffffu)),
                       deq((uint16_t)(u.y >> 16)));
}
// Density8 (u8): linear load (density8.hpp); the store needs the
// stochastic-rounding stream and is done in sweep_block16.
__device__ __forceinline__ float deq(uint8_t v) { return (float)v * (1.0f / 255.0f); }
__device__ __forceinline__ float4 load4(const uint8_t *p) {
    const uint32_t u = ld_global<uint32_t>(p);
    return make_float4(deq((uint8_t)(u & 0xffu)), deq((uint8_t)((u >> 8) & 0xffu)), deq((uint8_t)((u >> 16) & 0xffu)),
                       deq((uint8_t)(u >> 24)));
}
// The stochastic-rounding stream of a Density8 sweep: seed and a device
// counter of the sweeps done (read once per work-group).
struct SrCtx {
    uint32_t seed;
    const uint32_t *ctr;
};
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

// Row-wise fill of the dense gather tile (HSX^3, x fastest) for blocks whose
// rows lie inside the domain in x: a tile row (hy, hz) is resolved once
// under the boundary condition `bc`; its 16 interior cells are one block row
// (four float4 loads, or two float2 loads duplicated for a coarse neighbour)
// and its 2H x-halo cells are resolved and loaded one by one.  A row outside
// the domain (Dirichlet) is empty_f32.  `full == false` leaves the x-halo
// cells of rows outside the block in y or z at 0, as fill_tile_dense does.
template <int HALO, bool MULTI>
__device__ __forceinline__ void fill_tile_dense_rows(float *tile, const GridView &g, const GridView &c,
                                                     const int *s_nb, const int *s_nbc, int bx0, int by0, int bz0,
                                                     int bc, float empty_f32, bool full) {
    constexpr int B = 16, L = 4, HSX = B + 2 * HALO, n = 2 * HALO + 1, NTHR = kGatherThreads;
    constexpr int NR = HSX * HSX, NQ = 4 * NR, NH = 2 * HALO * NR;
    constexpr int QI = (NQ + NTHR - 1) / NTHR, HI = (NH + NTHR - 1) / NTHR;
    const int x0 = bx0 * B, y0 = by0 * B, z0 = bz0 * B;
    const int t = threadIdx.x;
    const float *qp[QI];
    uint32_t qcoarse = 0;
#pragma unroll
    for (int k = 0; k < QI; ++k) {  // interior quads: item = row * 4 + q
        const int it = t + k * NTHR;
        qp[k] = nullptr;
        if (it >= NQ) continue;
        const int r = it >> 2, q = it & 3;
        const int hy = r % HSX, hz = r / HSX;
        int ry, rz;
        if (!resolve_axis(y0 + hy - HALO, g.sy - 1, bc, ry) || !resolve_axis(z0 + hz - HALO, g.sz - 1, bc, rz))
            continue;
        const int nb = HALO + n * ((ry >> L) - by0 + HALO) + n * n * ((rz >> L) - bz0 + HALO);
        if (MULTI && s_nbc[nb] >= 0) {
            qcoarse |= 1u << k;
            qp[k] = c.slab + (int64_t)s_nbc[nb] * 512 + (((ry >> 1) & 7) << 3) + (((rz >> 1) & 7) << 6) + 2 * q;
        } else {
            qp[k] = g.slab + (int64_t)s_nb[nb] * 4096 + ((ry & 15) << 4) + ((rz & 15) << 8) + 4 * q;
        }
    }
    const float *hp[HI];
    uint32_t hhole = 0;
#pragma unroll
    for (int k = 0; k < HI; ++k) {  // x-halo cells: item = row * 2H + j; j < H left
        const int it = t + k * NTHR;
        hp[k] = nullptr;
        if (it >= NH) continue;
        const int r = it / (2 * HALO), j = it % (2 * HALO);
        const int hy = r % HSX, hz = r / HSX;
        const bool yh = hy < HALO || hy >= HALO + B, zh = hz < HALO || hz >= HALO + B;
        if (!full && (yh || zh)) {
            hhole |= 1u << k;
            continue;
        }
        const int gx = j < HALO ? x0 - HALO + j : x0 + B + (j - HALO);
        int rx, ry, rz;
        if (resolve_axis(gx, g.sx - 1, bc, rx) && resolve_axis(y0 + hy - HALO, g.sy - 1, bc, ry) &&
            resolve_axis(z0 + hz - HALO, g.sz - 1, bc, rz))
            hp[k] = cell_ptr<HALO, MULTI, float>(g.slab, c, s_nb, s_nbc, bx0, by0, bz0, rx, ry, rz);
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
            qv[k] = *reinterpret_cast<const float4 *>(qp[k]);
        }
    }
    float hv[HI];
#pragma unroll
    for (int k = 0; k < HI; ++k) hv[k] = hp[k] ? *hp[k] : (((hhole >> k) & 1u) ? 0.0f : empty_f32);
#pragma unroll
    for (int k = 0; k < QI; ++k) {
        const int it = t + k * NTHR;
        if (it >= NQ) continue;
        const int r = it >> 2, q = it & 3;
        float *d = tile + r * HSX + HALO + 4 * q;
        d[0] = qv[k].x;
        d[1] = qv[k].y;
        d[2] = qv[k].z;
        d[3] = qv[k].w;
    }
#pragma unroll
    for (int k = 0; k < HI; ++k) {
        const int it = t + k * NTHR;
        if (it >= NH) continue;
        const int r = it / (2 * HALO), j = it % (2 * HALO);
        tile[r * HSX + (j < HALO ? j : HALO + B + (j - HALO))] = hv[k];
    }
}

// Distance of tile coordinate h outside the block's [H, H + 16) (0 inside).
template <int HALO>
__device__ __forceinline__ int out_dist(int h) {
    return h < HALO ? HALO - h : (h >= HALO + 16 ? h - (HALO + 15) : 0);
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
// neighbour), its 2H x-halo cells are loaded one by one.  SKIP leaves out
// the tile cells the stencil never reads: kSkipCross (7-point: the corner
// rows and the x-halo of every row outside the block in y or z) or
// kSkipBiLap (25-tap: a row outside the block in y or z at distance 2 takes
// no x-halo, at distance 1 only the inner x-halo cell; of the corner rows only
// the four at distance 1 in both y and z, without x-halo).  An x-halo cell
// of a row outside the block lives in a diagonal neighbour block, so each
// skipped one saves a line fetched for a single value.
constexpr int kSkipNone = 0, kSkipCross = 1, kSkipBiLap = 2;

template <int HALO, bool MULTI, int NTHR, typename TS, int SKIP>
__device__ __forceinline__ void fill_tile_rows(float *tile, const TS *fs, const GridView &g, const GridView &c,
                                               const int *s_nb, const int *s_nbc, int bx0, int by0, int bz0,
                                               float empty_f32) {
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
        if (SKIP == kSkipCross && yh && zh) continue;
        if (SKIP == kSkipBiLap && yh && zh && (out_dist<HALO>(hy) != 1 || out_dist<HALO>(hz) != 1)) continue;
        qwrite |= 1u << k;
        int ry, rz;
        if (!resolve_axis(y0 + hy - HALO, g.sy - 1, MSBG_BC_NEUMANN, ry) ||
            !resolve_axis(z0 + hz - HALO, g.sz - 1, MSBG_BC_NEUMANN, rz))
            continue;
        const int nb = HALO + n * ((ry >> L) - by0 + HALO) + n * n * ((rz >> L) - bz0 + HALO);
        if (MULTI && s_nbc[nb] >= 0) {
            // the aligned half of the 8-cell coarse row that holds this
            // quad's two coarse cells (its elements 2 (q & 1), + 1)
            qcoarse |= 1u << k;
            qp[k] = reinterpret_cast<const TS *>(c.slab + (int64_t)s_nbc[nb] * 512 + (((ry >> 1) & 7) << 3) +
                                                 (((rz >> 1) & 7) << 6) + 4 * (q >> 1));
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
        if (SKIP == kSkipCross && (yh || zh)) continue;
        if (SKIP == kSkipBiLap && (yh || zh)) {
            // only rows at distance 1 outside (not corners) take the inner x-halo cell
            const bool inner = j == HALO - 1 || j == HALO;
            if ((yh && zh) || !inner || out_dist<HALO>(yh ? hy : hz) != 1) continue;
        }
        hwrite |= 1u << k;
        const int gx = j < HALO ? x0 - HALO + j : x0 + B + (j - HALO);
        int rx, ry, rz;
        if (resolve_axis(gx, g.sx - 1, MSBG_BC_NEUMANN, rx) &&
            resolve_axis(y0 + hy - HALO, g.sy - 1, MSBG_BC_NEUMANN, ry) &&
            resolve_axis(z0 + hz - HALO, g.sz - 1, MSBG_BC_NEUMANN, rz))
            hp[k] = cell_ptr<HALO, MULTI, TS>(fs, c, s_nb, s_nbc, bx0, by0, bz0, rx, ry, rz);
    }
    float4 qv[QI];
    float hv[HI];
    if constexpr (MULTI) {
        // One aligned 16-byte load per item, fine or coarse, none behind a
        // branch, so a lane's loads are all in flight together; an item with
        // no source reads the slab's first cells and takes the empty value.
        float4 raw[QI];
#pragma unroll
        for (int k = 0; k < QI; ++k) raw[k] = load4(qp[k] ? qp[k] : fs);
#pragma unroll
        for (int k = 0; k < HI; ++k) hv[k] = hp[k] ? deq(ld_global<TS>(hp[k])) : empty_f32;
#pragma unroll
        for (int k = 0; k < QI; ++k) {
            const int q = (t + k * NTHR) & 3;
            if (!qp[k])
                qv[k] = make_float4(empty_f32, empty_f32, empty_f32, empty_f32);
            else if ((qcoarse >> k) & 1u)
                qv[k] = (q & 1) ? make_float4(raw[k].z, raw[k].z, raw[k].w, raw[k].w)
                                : make_float4(raw[k].x, raw[k].x, raw[k].y, raw[k].y);
            else
                qv[k] = raw[k];
        }
    } else {
#pragma unroll
        for (int k = 0; k < QI; ++k)
            qv[k] = qp[k] ? load4(qp[k]) : make_float4(empty_f32, empty_f32, empty_f32, empty_f32);
#pragma unroll
        for (int k = 0; k < HI; ++k) hv[k] = hp[k] ? deq(ld_global<TS>(hp[k])) : empty_f32;
    }
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

// LDS floats of a 16^3 block's sweep tile (HSX^2 rows of kPitch).
template <int STENCIL>
constexpr int sweep_tile_floats() {
    return (16 + 2 * (STENCIL == MSBG_BILAPLACIAN ? 2 : 1)) * (16 + 2 * (STENCIL == MSBG_BILAPLACIAN ? 2 : 1)) *
           kPitch;
}

// The update of one 16^3 block `bid` whose topology row is `row`, by the
// whole work-group: tile, s_nb and s_nbc are its LDS.  The slabs come in as
// plain pointers taken from kernel arguments so every access compiles to a
// global (not flat) memory instruction.  512 lanes = 16 x 16 (x, y) x 2
// z-halves; each lane walks 8 cells up its z column keeping the planes it
// still needs in registers (5 / 9 / 13 LDS reads per cell for the 7 / 19 /
// 25-tap stencils once the window is primed).
template <int STENCIL, bool MULTI, typename TS>
__device__ __forceinline__ void sweep_block16(GridView g, GridView c, TS *__restrict__ fslab,
                                              const float *__restrict__ cslab, int64_t bid, const int32_t *row,
                                              float dt, int clamp01, float empty_f32, float *tile, int *s_nb,
                                              int *s_nbc, SrCtx sr) {
    static_assert(!MULTI || sizeof(TS) == sizeof(float), "cross-level sweeps are FP32 only");
    c.slab = const_cast<float *>(cslab);
    constexpr int HALO = STENCIL == MSBG_BILAPLACIAN ? 2 : 1;
    constexpr int HSX = 16 + 2 * HALO, NN = 2 * HALO + 1, NB = NN * NN * NN, CENTER = (NB - 1) / 2;
    constexpr int CPT = 16 * 256 / kSweepThreads;  // cells per lane along z
    const int slot = row[CENTER];
    if (slot < 2) return;  // dummy or absent: solver.rs:201-204 (uniform per workgroup)
    if (threadIdx.x < NB) {
        const int code = row[threadIdx.x];
        s_nb[threadIdx.x] = code >= 0 ? code : 0;
        if (MULTI) s_nbc[threadIdx.x] = code >= 0 ? -1 : -code - 1;
    }
    const int bx0 = (int)(bid % g.nx), by0 = (int)((bid / g.nx) % g.ny), bz0 = (int)(bid / ((int64_t)g.nx * g.ny));
    __syncthreads();
    if (bx0 * 16 + 16 <= g.sx)
        fill_tile_rows<HALO, MULTI, kSweepThreads, TS,
                       STENCIL == MSBG_LAPLACIAN ? kSkipCross
                                                 : (STENCIL == MSBG_BILAPLACIAN ? kSkipBiLap : kSkipNone)>(
            tile, fslab, g, c, s_nb, s_nbc, bx0, by0, bz0, empty_f32);
    else
        fill_tile_pitched_generic<HALO, MULTI, TS>(tile, fslab, g, c, s_nb, s_nbc, bx0, by0, bz0, empty_f32);
    __syncthreads();

    const int x = threadIdx.x & 15, y = (threadIdx.x >> 4) & 15, zs = (threadIdx.x >> 8) * CPT;
    TS *out = fslab + (int64_t)slot * 4096 + y * 16 + x;
    const bool cl = clamp01 != 0;
    uint32_t sr_key = 0;
    if constexpr (std::is_same<TS, uint8_t>::value) sr_key = sr_sweep_key(sr.seed, *sr.ctr);
    // StoreBack of cell (x, y, z)
    auto put = [&](int z, float v) {
        if constexpr (std::is_same<TS, uint8_t>::value)
            out[z * 256] = sr_quant8(v, sr_uniform(sr_key, (uint32_t)bid, (uint32_t)(z * 256 + y * 16 + x)));
        else
            store_back(out + z * 256, v);
    };
    // T(dx, dy, z): tile value at block cell (x+dx, y+dy, z).  The lane's
    // base is the low corner (x - H, y - H, zs - H) of its stencils, so every
    // read is the base plus a constant (z is zs plus a constant at each call).
    const float *tb = tile + tile_at<HALO>(x, y, zs);
    auto T = [&](int dx, int dy, int z) -> float {
        return tb[((z - zs + HALO) * HSX + (dy + HALO)) * kPitch + (dx + HALO)];
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
            put(z, cl ? clamp_f(v, 0.0f, 1.0f) : v);
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
            put(z, cl ? clamp_f(v, 0.0f, 1.0f) : v);
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
            put(z, cl ? clamp_f(v, 0.0f, 1.0f) : v);
            cm2 = pm.c;
            pm = p0;
            p0 = pp;
            pp = pp2;
        }
    }
}

// One colour pass: workgroup k updates block bids[k] in place, its halo
// topology in topo[k * (2H+1)^3 ...].
template <int STENCIL, bool MULTI, typename TS>
__global__ __launch_bounds__(kSweepThreads) void sweep_kernel(GridView g, GridView c, TS *__restrict__ fslab,
                                                              const float *__restrict__ cslab, const int64_t *bids,
                                                              const int32_t *topo, float dt, int clamp01,
                                                              float empty_f32, SrCtx sr) {
    constexpr int HALO = STENCIL == MSBG_BILAPLACIAN ? 2 : 1;
    constexpr int NN = 2 * HALO + 1, NB = NN * NN * NN;
    __shared__ __attribute__((aligned(16))) float tile[sweep_tile_floats<STENCIL>()];
    __shared__ int s_nb[NB];
    __shared__ int s_nbc[MULTI ? NB : 1];
    const int wg = xcd_remap(blockIdx.x, gridDim.x);
    sweep_block16<STENCIL, MULTI, TS>(g, c, fslab, cslab, bids[wg], topo + (int64_t)wg * NB, dt, clamp01, empty_f32,
                                      tile, s_nb, s_nbc, sr);
}

// Advances a Density8 plan's sweep counter (after the 8 colour passes).
__global__ void sr_advance_kernel(uint32_t *ctr) {
    if (threadIdx.x == 0) *ctr += 1u;
}

// ---------------------------------------------------------------------------
// Blocks at coarse levels.  No reference counterpart: the reference's
// fill_multires (halo.rs:187-358) fills a fine centre block only.  A block at
// level l >= 1 (edge E = 16 >> l) reads each halo voxel from the level its
// block lives at: its own level, the next coarser one (injection at g >> 1,
// as fill_multires reads a coarser neighbour, halo.rs:263-268) or the next
// finer one (the 2^3 mean of downsample_channel_avg, downsample.rs:44-53,
// summed dz, dy, dx in order).  Levels share one block grid.
// ---------------------------------------------------------------------------
constexpr int kMaxLevels = 5;

struct LevelSet {
    float *slab[kMaxLevels];
    int sx[kMaxLevels], sy[kMaxLevels], sz[kMaxLevels];
    int nx, ny, nz, n_levels;
};

// A work-group's share of one colour pass at a coarse level: the blocks
// cbids[first, first + count) of level `level`.
struct CoarseItem {
    int32_t level, first, count, pad;
};

// Topology code of each of a coarse block's 27 neighbours (index dx + 3 dy +
// 9 dz): slot | kind << 30 with kind 0 = the centre's level, 1 = the next
// coarser level, 2 = the next finer level.
constexpr uint32_t kKindShift = 30, kSlotMask = (1u << 30) - 1;

// Blocks per work-group at edge E and halo H: one cell per lane, the tiles
// within the 16^3 path's LDS tile.
__host__ __device__ constexpr int coarse_bpw(int E, int H, int tile_floats) {
    int b = kSweepThreads / (E * E * E);
    while (b > 1 && b * (E + 2 * H) * (E + 2 * H) * (E + 2 * H) > tile_floats) b >>= 1;
    return b;
}

// One cell's update from its tile taps T(dx, dy, dz), in the reference
// kernels' operation order (laplacian.rs:44-61, meancurv.rs:49-100,
// bilaplacian.rs:47-88) -- the same arithmetic as the 16^3 path.
template <int STENCIL, class Tap>
__device__ __forceinline__ float stencil_point(Tap T, float dt, bool cl) {
    float v;
    if (STENCIL == MSBG_LAPLACIAN) {
        const float cc = T(0, 0, 0), lx = T(-1, 0, 0), rx = T(1, 0, 0), dyu = T(0, -1, 0), dyd = T(0, 1, 0),
                    dzb = T(0, 0, -1), dzf = T(0, 0, 1);
        const float nb = lx + rx + dyu + dyd + dzb + dzf;
        const float lap = __builtin_fmaf(cc, -6.0f, nb);
        const float fx = (rx - lx) * 0.5f, fy = (dyd - dyu) * 0.5f, fz = (dzf - dzb) * 0.5f;
        const float gm = __builtin_sqrtf(__builtin_fmaf(fx, fx, __builtin_fmaf(fy, fy, fz * fz)));
        v = __builtin_fmaf(dt * lap, gm, cc);
    } else if (STENCIL == MSBG_MEAN_CURVATURE) {
        const float f0 = T(0, 0, 0), f1 = T(-1, 0, 0), f2 = T(1, 0, 0), f3 = T(0, -1, 0), f4 = T(0, 1, 0),
                    f5 = T(0, 0, -1), f6 = T(0, 0, 1);
        const float fx = (f2 - f1) * 0.5f, fy = (f4 - f3) * 0.5f, fz = (f6 - f5) * 0.5f;
        const float fxx = f1 + f2 - 2.0f * f0, fyy = f3 + f4 - 2.0f * f0, fzz = f5 + f6 - 2.0f * f0;
        const float fyx = (T(1, 1, 0) - T(1, -1, 0) - T(-1, 1, 0) + T(-1, -1, 0)) * 0.25f;
        const float fzx = (T(1, 0, 1) - T(1, 0, -1) - T(-1, 0, 1) + T(-1, 0, -1)) * 0.25f;
        const float fzy = (T(0, 1, 1) - T(0, 1, -1) - T(0, -1, 1) + T(0, -1, -1)) * 0.25f;
        const float crs = fx * (fy * fyx) + fx * (fz * fzx) + fy * (fz * fzy);
        const float hnum = __builtin_fmaf(fy * fy + fz * fz, fxx,
                                          __builtin_fmaf(fx * fx + fz * fz, fyy,
                                                         __builtin_fmaf(fx * fx + fy * fy, fzz, -2.0f * crs)));
        const float grad = __builtin_fmaf(fx, fx, __builtin_fmaf(fy, fy, fz * fz));
        const float hc = grad > 1e-7f ? hnum / grad : 0.0f;
        v = f0 + clamp_f(dt * hc, -0.1f, 0.1f);
    } else {
        const float f0 = T(0, 0, 0), f1 = T(-1, 0, 0), f2 = T(1, 0, 0), f3 = T(0, -1, 0), f4 = T(0, 1, 0),
                    f5 = T(0, 0, -1), f6 = T(0, 0, 1);
        const float s1 = f1 + f2 + f3 + f4 + f5 + f6;
        const float s2 = T(-2, 0, 0) + T(2, 0, 0) + T(0, -2, 0) + T(0, 2, 0) + T(0, 0, -2) + T(0, 0, 2);
        const float sd = T(1, 1, 0) + T(1, -1, 0) + T(-1, 1, 0) + T(-1, -1, 0) + T(1, 0, 1) + T(1, 0, -1) +
                         T(-1, 0, 1) + T(-1, 0, -1) + T(0, 1, 1) + T(0, 1, -1) + T(0, -1, 1) + T(0, -1, -1);
        const float hh = 42.0f * f0 - 12.0f * s1 + s2 + 2.0f * sd;
        const float fx = (f2 - f1) * 0.5f, fy = (f4 - f3) * 0.5f, fz = (f6 - f5) * 0.5f;
        const float gm = __builtin_sqrtf(__builtin_fmaf(fx, fx, __builtin_fmaf(fy, fy, fz * fz)));
        const float upd = (dt * hh) * gm;
        v = f0 + clamp_f(-upd, -0.1f, 0.1f);
    }
    return cl ? clamp_f(v, 0.0f, 1.0f) : v;
}

// One work item of coarse blocks: every lane fills tile cells (one dense
// (E + 2H)^3 tile per block), then updates one cell in place.
template <int STENCIL, int E>
__device__ __forceinline__ void sweep_coarse(const LevelSet &ls, const CoarseItem &it,
                                             const int32_t *__restrict__ cbids, const int32_t *__restrict__ ctopo,
                                             float dt, int clamp01, float *tile) {
    constexpr int H = STENCIL == MSBG_BILAPLACIAN ? 2 : 1;
    static_assert(H <= E, "a halo must stay within the 26 neighbour blocks");
    constexpr int LE = E == 8 ? 3 : (E == 4 ? 2 : (E == 2 ? 1 : 0));
    constexpr int T = E + 2 * H, T3 = T * T * T, E3 = E * E * E;
    const int l = it.level;
    const float *own = ls.slab[l];
    const float *crs = ls.slab[l + 1 < ls.n_levels ? l + 1 : l];
    const float *fin = ls.slab[l - 1];
    const int ntile = it.count * T3;
    for (int i = threadIdx.x; i < ntile; i += kSweepThreads) {
        const int j = i / T3, r = i - j * T3;
        const int hx = r % T, hy = (r / T) % T, hz = r / (T * T);
        const int bid = cbids[it.first + j];
        const int bx = bid % ls.nx, by = (bid / ls.nx) % ls.ny, bz = bid / (ls.nx * ls.ny);
        int rx, ry, rz;
        resolve_axis(bx * E + hx - H, ls.sx[l] - 1, MSBG_BC_NEUMANN, rx);
        resolve_axis(by * E + hy - H, ls.sy[l] - 1, MSBG_BC_NEUMANN, ry);
        resolve_axis(bz * E + hz - H, ls.sz[l] - 1, MSBG_BC_NEUMANN, rz);
        const int t = ((rx >> LE) - bx + 1) + 3 * ((ry >> LE) - by + 1) + 9 * ((rz >> LE) - bz + 1);
        const uint32_t code = (uint32_t)ctopo[(int64_t)(it.first + j) * 27 + t];
        const uint32_t kind = code >> kKindShift;
        const int64_t slot = code & kSlotMask;
        float v;
        if (kind == 0) {
            v = own[slot * E3 + ((rx & (E - 1)) | ((ry & (E - 1)) << LE) | ((rz & (E - 1)) << (2 * LE)))];
        } else if (kind == 1) {
            if constexpr (E >= 2) {
                constexpr int EC = E / 2, LC = LE - 1;
                v = crs[slot * (EC * EC * EC) + (((rx >> 1) & (EC - 1)) | (((ry >> 1) & (EC - 1)) << LC) |
                                                 (((rz >> 1) & (EC - 1)) << (2 * LC)))];
            } else {
                v = 0.0f;  // edge-1 blocks live at the coarsest level
            }
        } else {
            constexpr int EF = 2 * E, LF = LE + 1;
            const float *pf = fin + slot * (EF * EF * EF);
            float sum = 0.0f;
#pragma unroll
            for (int dz = 0; dz < 2; ++dz)
#pragma unroll
                for (int dy = 0; dy < 2; ++dy)
#pragma unroll
                    for (int dx = 0; dx < 2; ++dx)
                        sum += pf[((2 * rx + dx) & (EF - 1)) | (((2 * ry + dy) & (EF - 1)) << LF) |
                                  (((2 * rz + dz) & (EF - 1)) << (2 * LF))];
            v = sum / 8.0f;
        }
        tile[i] = v;
    }
    __syncthreads();
    const int j = threadIdx.x / E3;
    if (j >= it.count) return;
    const int cidx = threadIdx.x - j * E3;
    const int x = cidx & (E - 1), y = (cidx >> LE) & (E - 1), z = cidx >> (2 * LE);
    const float *tb = tile + j * T3;
    auto tap = [&](int dx, int dy, int dz) { return tb[((z + H + dz) * T + (y + H + dy)) * T + (x + H + dx)]; };
    const float v = stencil_point<STENCIL>(tap, dt, clamp01 != 0);
    const int64_t slot = (uint32_t)ctopo[(int64_t)(it.first + j) * 27 + 13] & kSlotMask;
    ls.slab[l][slot * E3 + cidx] = v;
}

// One colour pass over every level: work-groups [0, n0) update level-0
// blocks (bids0, topology rows topo0 in the single-level format with coarse
// codes), the rest take a CoarseItem each.  Blocks of one colour are never
// neighbours, whatever their levels, so they run in one launch.
template <int STENCIL>
__global__ __launch_bounds__(kSweepThreads) void sweep_levels_kernel(
    LevelSet ls, GridView g0, GridView g1, float *__restrict__ fslab, const float *__restrict__ cslab,
    const int64_t *__restrict__ bids0, const int32_t *__restrict__ topo0, int n0,
    const CoarseItem *__restrict__ items, const int32_t *__restrict__ cbids, const int32_t *__restrict__ ctopo,
    float dt, int clamp01) {
    constexpr int HALO = STENCIL == MSBG_BILAPLACIAN ? 2 : 1;
    constexpr int NN = 2 * HALO + 1, NB = NN * NN * NN;
    __shared__ __attribute__((aligned(16))) float tile[sweep_tile_floats<STENCIL>()];
    __shared__ int s_nb[NB];
    __shared__ int s_nbc[NB];
    const int wg = xcd_remap(blockIdx.x, gridDim.x);
    if (wg < n0) {
        sweep_block16<STENCIL, true, float>(g0, g1, fslab, cslab, bids0[wg], topo0 + (int64_t)wg * NB, dt, clamp01,
                                            0.0f, tile, s_nb, s_nbc, SrCtx{0u, nullptr});
        return;
    }
    const CoarseItem it = items[wg - n0];
    if (it.level == 1)
        sweep_coarse<STENCIL, 8>(ls, it, cbids, ctopo, dt, clamp01, tile);
    else if (it.level == 2)
        sweep_coarse<STENCIL, 4>(ls, it, cbids, ctopo, dt, clamp01, tile);
    else if (it.level == 3)
        sweep_coarse<STENCIL, 2>(ls, it, cbids, ctopo, dt, clamp01, tile);
    else if constexpr (HALO == 1)
        sweep_coarse<STENCIL, 1>(ls, it, cbids, ctopo, dt, clamp01, tile);
}

// Standalone halo gather: workgroup k writes the dense tile of bids[k].
template <int HALO, bool MULTI>
__global__ __launch_bounds__(kGatherThreads) void gather_kernel(GridView g, GridView c, const uint8_t *levels,
                                                                const int64_t *bids, int bc, int full,
                                                                float empty_f32, float *tiles) {
    constexpr int HSX = 16 + 2 * HALO, NT = HSX * HSX * HSX;
    __shared__ __attribute__((aligned(16))) float tile[NT];
    __shared__ int s_nb[125];
    __shared__ int s_nbc[MULTI ? 125 : 1];
    const int64_t bid = bids[blockIdx.x];
    const int bx0 = (int)(bid % g.nx), by0 = (int)((bid / g.nx) % g.ny), bz0 = (int)(bid / ((int64_t)g.nx * g.ny));
    resolve_neighbours<HALO, MULTI>(g, c, levels, bx0, by0, bz0, bid, s_nb, s_nbc);
    __syncthreads();
    if (bx0 * 16 + 16 <= g.sx)
        fill_tile_dense_rows<HALO, MULTI>(tile, g, c, s_nb, s_nbc, bx0, by0, bz0, bc, empty_f32, full != 0);
    else
        fill_tile_dense<HALO, MULTI>(tile, g, c, s_nb, s_nbc, bx0, by0, bz0, bc, empty_f32, full != 0);
    __syncthreads();
    // the tile is NT contiguous floats (NT and the tile stride are multiples of 4)
    static_assert(NT % 4 == 0, "tile rows are written as float4");
    float4 *dst = reinterpret_cast<float4 *>(tiles + (int64_t)blockIdx.x * NT);
    const float4 *src = reinterpret_cast<const float4 *>(tile);
    for (int i = threadIdx.x; i < NT / 4; i += kGatherThreads) dst[i] = src[i];
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
    uint8_t *slab8 = nullptr;    // Density8 (u8) storage over g's layout, or null
    uint32_t sr_seed = 0;        // Density8: stochastic-rounding seed
    uint32_t *d_sr_ctr = nullptr;  // Density8: sweeps done by this plan (device)
    // All-level plans (msbg_sweep_plan_new_levels): g / coarse are levels 0
    // and 1, d_bids / d_topo / counts the level-0 blocks; the coarse blocks
    // are d_cbids (colour-major, then level, then Morton order) with 27-entry
    // rows d_ctopo, grouped into work items per halo width.
    int n_levels = 0;
    msbg_grid *lv[kMaxLevels] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    int32_t *d_cbids = nullptr, *d_ctopo = nullptr;
    CoarseItem *d_items[3] = {nullptr, nullptr, nullptr};  // by halo width 1, 2
    int64_t item_counts[3][8] = {{0}};
    bool has_edge1 = false;  // blocks of edge 1 (level 4): radius-1 stencils only
    // Replayed sweeps (run_graph): one instantiated graph per (stencil, dt,
    // iterations, clamp01), least recently used evicted.  `done` is recorded
    // after every launch of an entry, so an evicted exec is destroyed only
    // once its last launch has finished.
    struct GraphEntry {
        hipGraphExec_t exec = nullptr;
        hipEvent_t done = nullptr;
        int stencil = -1, iters = -1, clamp = -1;
        float dt = 0.0f;
        uint64_t last_use = 0;
    };
    static constexpr int kGraphCache = 4;
    hipStream_t cap_stream = nullptr;
    GraphEntry graphs[kGraphCache];
    uint64_t graph_clock = 0;
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
                   const int32_t *topo, float dt, int clamp01, float empty_f32, SrCtx sr = SrCtx{0u, nullptr}) {
    const dim3 block(kSweepThreads);
    switch (stencil) {
        case MSBG_LAPLACIAN:
            hipLaunchKernelGGL((sweep_kernel<MSBG_LAPLACIAN, MULTI, TS>), grid, block, 0, s, g, c, fslab, c.slab, bids,
                               topo, dt, clamp01, empty_f32, sr);
            break;
        case MSBG_MEAN_CURVATURE:
            hipLaunchKernelGGL((sweep_kernel<MSBG_MEAN_CURVATURE, MULTI, TS>), grid, block, 0, s, g, c, fslab, c.slab,
                               bids, topo, dt, clamp01, empty_f32, sr);
            break;
        default:
            hipLaunchKernelGGL((sweep_kernel<MSBG_BILAPLACIAN, MULTI, TS>), grid, block, 0, s, g, c, fslab, c.slab,
                               bids, topo, dt, clamp01, empty_f32, sr);
    }
}

// Launches the 8 colour passes of one sweep on `s`.  When `ev` is non-null,
// an event pair brackets every launch.
template <bool MULTI>
int launch_sweep(msbg_sweep_plan *p, hipStream_t s, int stencil, float dt, int clamp01,
                 std::vector<hipEvent_t> *ev = nullptr) {
    GridView g = p->g->view();
    GridView c = MULTI ? p->coarse->view() : g;
    const float empty_f32 = MULTI ? 0.0f : (p->slab16 || p->slab8 ? 0.0f : p->g->empty_value);
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
            else if (!MULTI && p->slab8)
                launch_colour<false, uint8_t>(stencil, grid, s, g, c, p->slab8, bids, topo, dt, clamp01, empty_f32,
                                              SrCtx{p->sr_seed, p->d_sr_ctr});
            else
                launch_colour<MULTI, float>(stencil, grid, s, g, c, g.slab, bids, topo, dt, clamp01, empty_f32);
            MSBG_HIP(hipGetLastError());
            if (ev) MSBG_HIP(hipEventRecord(ev->back(), s));
        }
        off += cnt;
    }
    if (!MULTI && p->slab8) {
        hipLaunchKernelGGL(sr_advance_kernel, dim3(1), dim3(64), 0, s, p->d_sr_ctr);
        MSBG_HIP(hipGetLastError());
    }
    return MSBG_OK;
}

// The 8 colour passes of an all-level plan: one launch per colour covers
// the level-0 blocks and the coarse work items of that colour.
int launch_sweep_levels(msbg_sweep_plan *p, hipStream_t s, int stencil, float dt, int clamp01,
                        std::vector<hipEvent_t> *ev = nullptr) {
    const int halo = stencil == MSBG_BILAPLACIAN ? 2 : 1;
    const int nb = halo == 2 ? 125 : 27;
    LevelSet ls{};
    for (int l = 0; l < p->n_levels; ++l) {
        ls.slab[l] = p->lv[l]->d_slab;
        ls.sx[l] = (int)p->lv[l]->sx;
        ls.sy[l] = (int)p->lv[l]->sy;
        ls.sz[l] = (int)p->lv[l]->sz;
    }
    ls.nx = (int)p->g->nx;
    ls.ny = (int)p->g->ny;
    ls.nz = (int)p->g->nz;
    ls.n_levels = p->n_levels;
    GridView g0 = p->lv[0]->view(), g1 = p->lv[1]->view();
    int64_t off = 0, ioff = 0;
    for (int col = 0; col < 8; ++col) {
        const int64_t n0 = p->counts[col], ni = p->item_counts[halo][col];
        if (n0 + ni > 0) {
            if (ev) {
                hipEvent_t a, b;
                MSBG_HIP(hipEventCreate(&a));
                MSBG_HIP(hipEventCreate(&b));
                ev->push_back(a);
                ev->push_back(b);
                MSBG_HIP(hipEventRecord(a, s));
            }
            const dim3 grid((unsigned)(n0 + ni)), block(kSweepThreads);
            const int64_t *bids0 = p->d_bids + off;
            const int32_t *topo0 = p->d_topo[halo] + off * nb;
            const CoarseItem *items = p->d_items[halo] + ioff;
#define MSBG_LEVELS_LAUNCH(ST)                                                                                  \
    hipLaunchKernelGGL(sweep_levels_kernel<ST>, grid, block, 0, s, ls, g0, g1, g0.slab, g1.slab, bids0, topo0, \
                       (int)n0, items, p->d_cbids, p->d_ctopo, dt, clamp01)
            if (stencil == MSBG_LAPLACIAN)
                MSBG_LEVELS_LAUNCH(MSBG_LAPLACIAN);
            else if (stencil == MSBG_MEAN_CURVATURE)
                MSBG_LEVELS_LAUNCH(MSBG_MEAN_CURVATURE);
            else
                MSBG_LEVELS_LAUNCH(MSBG_BILAPLACIAN);
#undef MSBG_LEVELS_LAUNCH
            MSBG_HIP(hipGetLastError());
            if (ev) MSBG_HIP(hipEventRecord(ev->back(), s));
        }
        off += n0;
        ioff += ni;
    }
    return MSBG_OK;
}

int launch_iterations(msbg_sweep_plan *p, hipStream_t s, int stencil, float dt, int iterations, int clamp01,
                      std::vector<hipEvent_t> *ev = nullptr) {
    MSBG_CHECK(!(p->has_edge1 && stencil == MSBG_BILAPLACIAN), MSBG_EINVAL,
               "the radius-2 bi-Laplacian needs blocks of edge >= 2 (a level-4 halo would reach past its "
               "neighbour blocks)");
    for (int it = 0; it < iterations; ++it) {
        int rc = p->n_levels ? launch_sweep_levels(p, s, stencil, dt, clamp01, ev)
                 : p->coarse ? launch_sweep<true>(p, s, stencil, dt, clamp01, ev)
                             : launch_sweep<false>(p, s, stencil, dt, clamp01, ev);
        if (rc) return rc;
    }
    return MSBG_OK;
}

// An all-level plan: the level-0 blocks as a multires plan over levels 0 / 1,
// the coarse blocks with host-built 27-entry topology rows.
int plan_new_levels(msbg_grid *const *grids, int n_levels, const uint8_t *d_levels, const int64_t *active,
                    int64_t n, msbg_sweep_plan **out) {
    MSBG_CHECK(out, MSBG_EINVAL, "null out");
    *out = nullptr;
    MSBG_CHECK(grids && n_levels >= 2 && n_levels <= kMaxLevels, MSBG_EINVAL,
               "an all-level sweep needs 2 to 5 levels");
    MSBG_CHECK(d_levels, MSBG_EINVAL, "all-level sweep: null level map");
    MSBG_CHECK(n >= 0 && (n == 0 || active), MSBG_EINVAL, "all-level sweep: bad active list");
    for (int l = 0; l < n_levels; ++l) {
        MSBG_CHECK(grids[l], MSBG_EINVAL, "all-level sweep: null level grid");
        MSBG_CHECK(grids[l]->L == 4 - l && grids[l]->nx == grids[0]->nx && grids[l]->ny == grids[0]->ny &&
                       grids[l]->nz == grids[0]->nz,
                   MSBG_EINVAL, "all-level sweep: level l must have block edge 16 >> l over one block grid");
    }
    const msbg_grid *g0 = grids[0];
    MSBG_CHECK(g0->n_blocks < ((int64_t)1 << 30), MSBG_EINVAL, "all-level sweep: too many blocks");
    std::vector<uint8_t> lev((size_t)g0->n_blocks);
    MSBG_HIP(hipMemcpy(lev.data(), d_levels, lev.size(), hipMemcpyDeviceToHost));
    std::vector<int64_t> fine;
    std::vector<std::pair<uint64_t, int32_t>> coarse[8][kMaxLevels];  // (Morton key, bid)
    for (int64_t i = 0; i < n; ++i) {
        const int64_t bid = active[i];
        MSBG_CHECK(bid >= 0 && bid < g0->n_blocks, MSBG_EINVAL, "all-level sweep: bid out of range");
        const int l = lev[(size_t)bid];
        MSBG_CHECK(l < n_levels, MSBG_EINVAL, "all-level sweep: level map names a level beyond n_levels");
        if (l == 0) {
            fine.push_back(bid);
            continue;
        }
        if (grids[l]->b2s_host[(size_t)bid] < 2) continue;  // no value block at its level: skipped
        const uint64_t bx = bid % g0->nx, by = (bid / g0->nx) % g0->ny, bz = bid / g0->nxy;
        const int color = (int)((bx & 1) | ((by & 1) << 1) | ((bz & 1) << 2));
        coarse[color][l].push_back({spread21(bx) | (spread21(by) << 1) | (spread21(bz) << 2), (int32_t)bid});
    }
    msbg_sweep_plan *p = nullptr;
    int rc = plan_new(grids[0], grids[1], d_levels, fine.data(), (int64_t)fine.size(), &p);
    if (rc) return rc;
    p->n_levels = n_levels;
    for (int l = 0; l < n_levels; ++l) p->lv[l] = grids[l];
    std::vector<int32_t> cb, ct;
    std::vector<CoarseItem> items[3];
    for (int col = 0; col < 8; ++col)
        for (int l = 1; l < n_levels; ++l) {
            auto &v = coarse[col][l];
            std::sort(v.begin(), v.end());
            const int E = 16 >> l;
            if (E == 1 && !v.empty()) p->has_edge1 = true;
            const int first = (int)cb.size();
            for (auto &kb : v) {
                const int64_t bid = kb.second;
                cb.push_back((int32_t)bid);
                const int64_t bx = bid % g0->nx, by = (bid / g0->nx) % g0->ny, bz = bid / g0->nxy;
                for (int t = 0; t < 27; ++t) {
                    const int64_t x = bx + t % 3 - 1, y = by + (t / 3) % 3 - 1, z = bz + t / 9 - 1;
                    uint32_t code = 0;  // outside the block grid: never read (the mirror stays inside)
                    if (x >= 0 && x < g0->nx && y >= 0 && y < g0->ny && z >= 0 && z < g0->nz) {
                        const int64_t b = x + y * g0->nx + z * g0->nxy;
                        const int ln = lev[(size_t)b];
                        if (ln > l && l + 1 < n_levels)
                            code = (uint32_t)grids[l + 1]->b2s_host[(size_t)b] | (1u << kKindShift);
                        else if (ln < l)
                            code = (uint32_t)grids[l - 1]->b2s_host[(size_t)b] | (2u << kKindShift);
                        else
                            code = (uint32_t)grids[l]->b2s_host[(size_t)b];
                    }
                    ct.push_back((int32_t)code);
                }
            }
            for (int h = 1; h <= 2; ++h) {
                if (E < h) continue;
                const int bpw = coarse_bpw(E, h, h == 2 ? sweep_tile_floats<MSBG_BILAPLACIAN>()
                                                        : sweep_tile_floats<MSBG_LAPLACIAN>());
                for (int k = 0; k < (int)v.size(); k += bpw) {
                    items[h].push_back(CoarseItem{l, first + k, std::min(bpw, (int)v.size() - k), 0});
                    p->item_counts[h][col] += 1;
                }
            }
        }
    hipError_t e = hipSuccess;
    if (!cb.empty()) {
        e = hipMalloc(&p->d_cbids, cb.size() * sizeof(int32_t));
        if (e == hipSuccess) e = hipMalloc(&p->d_ctopo, ct.size() * sizeof(int32_t));
        if (e == hipSuccess) e = hipMemcpy(p->d_cbids, cb.data(), cb.size() * sizeof(int32_t), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(p->d_ctopo, ct.data(), ct.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    }
    for (int h = 1; h <= 2 && e == hipSuccess; ++h)
        if (!items[h].empty()) {
            e = hipMalloc(&p->d_items[h], items[h].size() * sizeof(CoarseItem));
            if (e == hipSuccess)
                e = hipMemcpy(p->d_items[h], items[h].data(), items[h].size() * sizeof(CoarseItem),
                              hipMemcpyHostToDevice);
        }
    if (e != hipSuccess) {
        msbg_sweep_plan_free(p);
        return hip_fail(e, "all-level sweep plan upload");
    }
    p->n += (int64_t)cb.size();
    *out = p;
    return MSBG_OK;
}

// Destroys a cached graph after its last launch has completed.
void drop_graph(msbg_sweep_plan::GraphEntry &e) {
    if (e.done) (void)hipEventSynchronize(e.done);
    if (e.exec) (void)hipGraphExecDestroy(e.exec);
    e.exec = nullptr;
    e.stencil = -1;
}

void drop_graphs(msbg_sweep_plan *p) {
    for (auto &e : p->graphs) {
        drop_graph(e);
        if (e.done) (void)hipEventDestroy(e.done);
        e.done = nullptr;
    }
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

int msbg_sweep_plan_new_density8(msbg_grid *layout, uint8_t *d_slab8, const int64_t *active, int64_t n,
                                 uint32_t seed, msbg_sweep_plan **out) {
    MSBG_CHECK(d_slab8, MSBG_EINVAL, "density8 plan: null u8 slab");
    int rc = plan_new(layout, nullptr, nullptr, active, n, out);
    if (rc) return rc;
    msbg_sweep_plan *p = *out;
    p->slab8 = d_slab8;
    p->sr_seed = seed;
    const uint32_t zero = 0;
    hipError_t e = hipMalloc(&p->d_sr_ctr, sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemcpy(p->d_sr_ctr, &zero, sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        msbg_sweep_plan_free(p);
        *out = nullptr;
        return hip_fail(e, "density8 plan");
    }
    return MSBG_OK;
}

int msbg_sweep_density8(msbg_grid *layout, uint8_t *d_slab8, const int64_t *active, int64_t n, int stencil, float dt,
                        int iterations, int clamp01, uint32_t seed) {
    int rc = check_stencil(stencil);
    if (rc) return rc;
    msbg_sweep_plan *p = nullptr;
    rc = msbg_sweep_plan_new_density8(layout, d_slab8, active, n, seed, &p);
    if (rc) return rc;
    rc = msbg_sweep_plan_run_async(p, stencil, dt, iterations, clamp01);
    if (rc == MSBG_OK) {
        hipError_t e = hipStreamSynchronize(stream());
        if (e != hipSuccess) rc = hip_fail(e, "msbg_sweep_density8");
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
    drop_graphs(p);
    if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
    if (p->d_bids) (void)hipFree(p->d_bids);
    for (int32_t *t : p->d_topo)
        if (t) (void)hipFree(t);
    if (p->d_sr_ctr) (void)hipFree(p->d_sr_ctr);
    if (p->d_cbids) (void)hipFree(p->d_cbids);
    if (p->d_ctopo) (void)hipFree(p->d_ctopo);
    for (CoarseItem *it : p->d_items)
        if (it) (void)hipFree(it);
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
    // a hit, else an empty entry, else the least recently used one
    msbg_sweep_plan::GraphEntry *hit = nullptr, *lru = nullptr;
    for (auto &e : p->graphs) {
        if (e.exec && e.stencil == stencil && e.iters == iterations && e.clamp == clamp01 && e.dt == dt) hit = &e;
        if (!lru || (lru->exec && (!e.exec || e.last_use < lru->last_use))) lru = &e;
    }
    if (!hit) {
        hit = lru;
        drop_graph(*hit);
        if (!hit->done) MSBG_HIP(hipEventCreateWithFlags(&hit->done, hipEventDisableTiming));
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
        e = hipGraphInstantiate(&hit->exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) hit->exec = nullptr;
        MSBG_HIP(e);
        hit->stencil = stencil;
        hit->iters = iterations;
        hit->clamp = clamp01;
        hit->dt = dt;
    }
    hit->last_use = ++p->graph_clock;
    MSBG_HIP(hipGraphLaunch(hit->exec, stream()));
    MSBG_HIP(hipEventRecord(hit->done, stream()));
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

int msbg_sweep_plan_new_levels(msbg_grid *const *grids, int n_levels, const uint8_t *dev_levels,
                               const int64_t *active, int64_t n, msbg_sweep_plan **out) {
    return plan_new_levels(grids, n_levels, dev_levels, active, n, out);
}

int msbg_sweep_levels(msbg_grid *const *grids, int n_levels, const uint8_t *dev_levels, const int64_t *active,
                      int64_t n, int stencil, float dt, int iterations, int clamp01) {
    int rc = check_stencil(stencil);
    if (rc) return rc;
    msbg_sweep_plan *p = nullptr;
    rc = plan_new_levels(grids, n_levels, dev_levels, active, n, &p);
    if (rc) return rc;
    rc = msbg_sweep_plan_run_async(p, stencil, dt, iterations, clamp01);
    if (rc == MSBG_OK) {
        hipError_t e = hipStreamSynchronize(stream());
        if (e != hipSuccess) rc = hip_fail(e, "msbg_sweep_levels");
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

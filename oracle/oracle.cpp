// CPU oracle for the sparse-block-grid smoother and particle-transfer path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, function by function, the
// algorithms of the msbg-rs reference (Rust, CPU) in plain C++ so that the HIP
// product can be checked against it.  It is loaded only by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as the
// checker or as the timed CPU baseline -- never as the product.
//
// Every function cites the reference file:line it follows.  Floating point
// follows the reference operation by operation: Rust's `a.mul_add(b, c)` is
// `fmaf(a, b, c)` here, every other expression is evaluated left to right
// without contraction (the file is compiled with -ffp-contract=off).
//
// Shared data layout (identical to the HIP product):
//   * a level with block edge B = 2^L has nx = ceil(sx/B) ... blocks,
//     bid = bx + by*nx + bz*nx*ny (sparse_grid.rs:159-165),
//     vid = vx | vy<<L | vz<<2L (sparse_grid.rs:169-175);
//   * `b2s[bid]` (int32) names the slab slot of a block: 0 = the empty dummy
//     (absent block), 1 = the full dummy, >= 2 = an allocated value block;
//   * `slab[slot * B^3 + vid]` holds the payload; slot 0 is filled with the
//     empty value and slot 1 with the full value (sparse_grid.rs:120-129).
//
// Parity status of each function is listed in the test files that pin it.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>
#include <xmmintrin.h>
#include <pmmintrin.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef int64_t i64;
typedef int32_t i32;
typedef uint64_t u64;
typedef uint32_t u32;

namespace {

struct Level {
    i64 sx, sy, sz, nx, ny, nz, nxy, nblocks;
    int L;       // log2 block edge
    int B;       // block edge
    i64 N;       // voxels per block
    const float *slab;
    const i32 *b2s;
    Level(i64 sx_, i64 sy_, i64 sz_, int L_, const float *slab_, const i32 *b2s_)
        : sx(sx_), sy(sy_), sz(sz_), L(L_), B(1 << L_), slab(slab_), b2s(b2s_) {
        nx = (sx + B - 1) >> L;
        ny = (sy + B - 1) >> L;
        nz = (sz + B - 1) >> L;
        nxy = nx * ny;
        nblocks = nxy * nz;
        N = (i64)B * B * B;
    }
    // SparseGrid::block_data_ptr (sparse_grid.rs:405-411): absent -> empty dummy.
    const float *block_ptr(i64 bid) const { return slab + (i64)b2s[bid] * N; }
    float empty_value() const { return slab[0]; }
};

enum Bc { BC_CLAMP = 0, BC_NEUMANN = 1, BC_DIRICHLET = 2 };

// boundary.rs:48-56 -- symmetric mirror with period 2(max+1).
inline i32 reflect(i32 g, i32 max) {
    i32 period = 2 * (max + 1);
    i32 r = g % period;
    if (r < 0) r += period;  // rem_euclid
    return r > max ? period - 1 - r : r;
}

// boundary.rs:35-44 -- returns false only for Dirichlet out of range.
inline bool resolve_axis(i32 g, i32 max, int bc, i32 *out) {
    if (g >= 0 && g <= max) { *out = g; return true; }
    if (bc == BC_CLAMP) { *out = g < 0 ? 0 : max; return true; }
    if (bc == BC_NEUMANN) { *out = reflect(g, max); return true; }
    return false;
}

// multires/sort.rs:8-16
inline u64 spread(u64 v) {
    v &= 0x1fffffULL;
    v = (v | (v << 32)) & 0x001f00000000ffffULL;
    v = (v | (v << 16)) & 0x001f0000ff0000ffULL;
    v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
    v = (v | (v << 2)) & 0x1249249249249249ULL;
    return v;
}

// Rust `simd_clamp(lo, hi)` = max(lo) then min(hi).
inline float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// A1: morton3 -- multires/sort.rs:20-22
// ---------------------------------------------------------------------------
u64 orc_morton3(u64 x, u64 y, u64 z) { return spread(x) | (spread(y) << 1) | (spread(z) << 2); }

// ---------------------------------------------------------------------------
// A4: build_color_buckets -- solver.rs:247-267.  Writes the 8 buckets back to
// back into `out` (n entries) and their sizes into counts[8].
// ---------------------------------------------------------------------------
void orc_color_buckets(i64 sx, i64 sy, i64 sz, int L, const i64 *active, i64 n, i64 *out, i64 *counts) {
    Level g(sx, sy, sz, L, nullptr, nullptr);
    std::vector<i64> buckets[8];
    for (i64 i = 0; i < n; ++i) {
        i64 bid = active[i];
        i64 bx = bid % g.nx, by = (bid / g.nx) % g.ny, bz = bid / g.nxy;
        int color = (int)((bx & 1) | ((by & 1) << 1) | ((bz & 1) << 2));
        buckets[color].push_back(bid);
    }
    i64 k = 0;
    for (int c = 0; c < 8; ++c) {
        auto key = [&](i64 bid) {
            i64 bx = bid % g.nx, by = (bid / g.nx) % g.ny, bz = bid / g.nxy;
            return orc_morton3((u64)bx, (u64)by, (u64)bz);
        };
        std::sort(buckets[c].begin(), buckets[c].end(), [&](i64 a, i64 b) { return key(a) < key(b); });
        counts[c] = (i64)buckets[c].size();
        for (i64 bid : buckets[c]) out[k++] = bid;
    }
}

// ---------------------------------------------------------------------------
// A6: HaloBlock::fill::<HALO, FULL, f32> -- multires/halo.rs:31-170.
// `tile` (HSX^3, HSX = B + 2*halo) is zeroed first, as a fresh HaloBlock
// (halo.rs:16-20); faces-only fills leave x-halo edge/corner cells at zero.
// ---------------------------------------------------------------------------
void orc_halo_fill(i64 sx, i64 sy, i64 sz, int L, const float *slab, const i32 *b2s, i64 bid, int halo,
                   int full, int bc, float *tile) {
    Level g(sx, sy, sz, L, slab, b2s);
    const int B = g.B;
    const int HSX = B + 2 * halo;
    std::memset(tile, 0, sizeof(float) * (size_t)HSX * HSX * HSX);
    const float empty_f32 = g.empty_value();
    const i32 bx0 = (i32)(bid % g.nx), by0 = (i32)((bid / g.nx) % g.ny), bz0 = (i32)(bid / g.nxy);
    const i32 x0 = bx0 * B;
    const int n = 2 * halo + 1;
    const float *ptrs[125];
    // halo.rs:59-80: resolve the (2H+1)^3 neighbourhood once.
    for (int dz = 0; dz < n; ++dz)
        for (int dy = 0; dy < n; ++dy)
            for (int dx = 0; dx < n; ++dx) {
                i32 bx = bx0 + dx - halo, by = by0 + dy - halo, bz = bz0 + dz - halo;
                const float *p = g.slab;  // empty dummy
                if (bx >= 0 && bx < g.nx && by >= 0 && by < g.ny && bz >= 0 && bz < g.nz)
                    p = g.block_ptr((i64)bx + (i64)by * g.nx + (i64)bz * g.nxy);
                ptrs[dx + n * dy + n * n * dz] = p;
            }
    const i32 mask = B - 1;
    auto read = [&](i32 x, i32 y, i32 z) -> float {  // halo.rs:87-96
        int idx = ((x >> L) - bx0 + halo) + n * ((y >> L) - by0 + halo) + n * n * ((z >> L) - bz0 + halo);
        i64 vid = (x & mask) | ((y & mask) << L) | ((i64)(z & mask) << (2 * L));
        return ptrs[idx][vid];
    };
    i32 rxl[2], rxr[2];
    bool okl[2], okr[2];
    for (int j = 0; j < halo; ++j) {  // halo.rs:99-104
        okl[j] = resolve_axis(x0 - halo + j, (i32)sx - 1, bc, &rxl[j]);
        okr[j] = resolve_axis(x0 + B + j, (i32)sx - 1, bc, &rxr[j]);
    }
    for (int z = 0; z < HSX; ++z) {
        i32 rz = 0;
        bool okz = resolve_axis(bz0 * B + z - halo, (i32)sz - 1, bc, &rz);
        for (int y = 0; y < HSX; ++y) {
            i32 ry = 0;
            bool oky = resolve_axis(by0 * B + y - halo, (i32)sy - 1, bc, &ry);
            float *row = tile + (size_t)z * HSX * HSX + (size_t)y * HSX;
            if (oky && okz && x0 + B <= sx) {  // halo.rs:117-129 contiguous middle row
                int idx = halo + n * ((ry >> L) - by0 + halo) + n * n * ((rz >> L) - bz0 + halo);
                i64 vid = ((i64)(ry & mask) << L) | ((i64)(rz & mask) << (2 * L));
                std::memcpy(row + halo, ptrs[idx] + vid, sizeof(float) * B);
            } else if (oky && okz) {  // halo.rs:130-144 partial last block
                for (int i = 0; i < B; ++i) {
                    i32 rx;
                    row[halo + i] = resolve_axis(x0 + i, (i32)sx - 1, bc, &rx) ? read(rx, ry, rz) : empty_f32;
                }
            } else {
                for (int i = 0; i < B; ++i) row[halo + i] = empty_f32;
            }
            if (full || (y >= halo && y < halo + B && z >= halo && z < halo + B)) {  // halo.rs:154-167
                for (int j = 0; j < halo; ++j) {
                    row[j] = (okl[j] && oky && okz) ? read(rxl[j], ry, rz) : empty_f32;
                    row[halo + B + j] = (okr[j] && oky && okz) ? read(rxr[j], ry, rz) : empty_f32;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// A7: fill_multires -- multires/halo.rs:187-358.  `fine` is level l (block
// edge 2^L), `coarse` is level l+1 (edge 2^(L-1)), both over one block grid.
// A neighbour b is coarse iff levels[b] > levels[bid] (halo.rs:243-247); a
// coarse read samples (x>>1, y>>1, z>>1) inside coarse block b (halo.rs:263-268).
// Empty reads are 0.0 (halo.rs:215).
// ---------------------------------------------------------------------------
void orc_halo_fill_multires(i64 sx, i64 sy, i64 sz, int L, const float *fine_slab, const i32 *fine_b2s,
                            const float *coarse_slab, const i32 *coarse_b2s, const uint8_t *levels, i64 bid,
                            int halo, int full, int bc, float *tile) {
    Level f(sx, sy, sz, L, fine_slab, fine_b2s);
    const int B = f.B;
    const int Lc = L - 1;
    const i64 Nc = (i64)1 << (3 * Lc);
    const int HSX = B + 2 * halo;
    std::memset(tile, 0, sizeof(float) * (size_t)HSX * HSX * HSX);
    const float empty_f32 = 0.0f;
    const i32 bx0 = (i32)(bid % f.nx), by0 = (i32)((bid / f.nx) % f.ny), bz0 = (i32)(bid / f.nxy);
    const i32 x0 = bx0 * B;
    const uint8_t lvl = levels[bid];
    const int n = 2 * halo + 1;
    const float *ptrs[125];
    const float *cptrs[125];
    for (int dz = 0; dz < n; ++dz)
        for (int dy = 0; dy < n; ++dy)
            for (int dx = 0; dx < n; ++dx) {
                int idx = dx + n * dy + n * n * dz;
                ptrs[idx] = f.slab;
                cptrs[idx] = nullptr;
                i32 bx = bx0 + dx - halo, by = by0 + dy - halo, bz = bz0 + dz - halo;
                if (bx >= 0 && bx < f.nx && by >= 0 && by < f.ny && bz >= 0 && bz < f.nz) {
                    i64 b = (i64)bx + (i64)by * f.nx + (i64)bz * f.nxy;
                    if (levels[b] > lvl)
                        cptrs[idx] = coarse_slab + (i64)coarse_b2s[b] * Nc;
                    else
                        ptrs[idx] = f.block_ptr(b);
                }
            }
    const i32 mask = B - 1, cmask = (B >> 1) - 1;
    auto read = [&](i32 x, i32 y, i32 z) -> float {  // halo.rs:258-275
        int idx = ((x >> L) - bx0 + halo) + n * ((y >> L) - by0 + halo) + n * n * ((z >> L) - bz0 + halo);
        const float *cp = cptrs[idx];
        if (cp) {
            i64 vid = ((x >> 1) & cmask) | (((i64)(y >> 1) & cmask) << Lc) | (((i64)(z >> 1) & cmask) << (2 * Lc));
            return cp[vid];
        }
        i64 vid = (x & mask) | ((i64)(y & mask) << L) | ((i64)(z & mask) << (2 * L));
        return ptrs[idx][vid];
    };
    i32 rxl[2], rxr[2];
    bool okl[2], okr[2];
    for (int j = 0; j < halo; ++j) {
        okl[j] = resolve_axis(x0 - halo + j, (i32)sx - 1, bc, &rxl[j]);
        okr[j] = resolve_axis(x0 + B + j, (i32)sx - 1, bc, &rxr[j]);
    }
    for (int z = 0; z < HSX; ++z) {
        i32 rz = 0;
        bool okz = resolve_axis(bz0 * B + z - halo, (i32)sz - 1, bc, &rz);
        for (int y = 0; y < HSX; ++y) {
            i32 ry = 0;
            bool oky = resolve_axis(by0 * B + y - halo, (i32)sy - 1, bc, &ry);
            float *row = tile + (size_t)z * HSX * HSX + (size_t)y * HSX;
            if (oky && okz && x0 + B <= sx) {
                int idx = halo + n * ((ry >> L) - by0 + halo) + n * n * ((rz >> L) - bz0 + halo);
                if (cptrs[idx]) {  // halo.rs:299-311
                    for (int i = 0; i < B; ++i) {
                        i32 rx;
                        row[halo + i] = resolve_axis(x0 + i, (i32)sx - 1, bc, &rx) ? read(rx, ry, rz) : empty_f32;
                    }
                } else {
                    i64 vid = ((i64)(ry & mask) << L) | ((i64)(rz & mask) << (2 * L));
                    std::memcpy(row + halo, ptrs[idx] + vid, sizeof(float) * B);
                }
            } else if (oky && okz) {
                for (int i = 0; i < B; ++i) {
                    i32 rx;
                    row[halo + i] = resolve_axis(x0 + i, (i32)sx - 1, bc, &rx) ? read(rx, ry, rz) : empty_f32;
                }
            } else {
                for (int i = 0; i < B; ++i) row[halo + i] = empty_f32;
            }
            if (full || (y >= halo && y < halo + B && z >= halo && z < halo + B)) {
                for (int j = 0; j < halo; ++j) {
                    row[j] = (okl[j] && oky && okz) ? read(rxl[j], ry, rz) : empty_f32;
                    row[halo + B + j] = (okr[j] && oky && okz) ? read(rxr[j], ry, rz) : empty_f32;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Halo topology: the neighbour pointers HaloBlock::fill resolves
// (halo.rs:59-80) and fill_multires resolves (halo.rs:234-251), written as
// slab slots.  Entry dx + n*dy + n*n*dz: the neighbour's fine slot (0 = empty
// dummy, also for blocks outside the block grid), or -(coarse slot) - 1 when
// levels[b] > levels[bid].  coarse_b2s / levels may be null (single level).
// ---------------------------------------------------------------------------
void orc_halo_topology(i64 sx, i64 sy, i64 sz, int L, const i32 *b2s, const i32 *coarse_b2s,
                       const uint8_t *levels, i64 bid, int halo, i32 *out) {
    Level g(sx, sy, sz, L, nullptr, b2s);
    const i32 bx0 = (i32)(bid % g.nx), by0 = (i32)((bid / g.nx) % g.ny), bz0 = (i32)(bid / g.nxy);
    const int n = 2 * halo + 1;
    for (int dz = 0; dz < n; ++dz)
        for (int dy = 0; dy < n; ++dy)
            for (int dx = 0; dx < n; ++dx) {
                i32 bx = bx0 + dx - halo, by = by0 + dy - halo, bz = bz0 + dz - halo;
                i32 code = 0;
                if (bx >= 0 && bx < g.nx && by >= 0 && by < g.ny && bz >= 0 && bz < g.nz) {
                    i64 b = (i64)bx + (i64)by * g.nx + (i64)bz * g.nxy;
                    if (levels && coarse_b2s && levels[b] > levels[bid])
                        code = -coarse_b2s[b] - 1;
                    else
                        code = b2s[b];
                }
                out[dx + n * dy + n * n * dz] = code;
            }
}

// ---------------------------------------------------------------------------
// A8: kernel_laplacian -- math/laplacian.rs:16-70 (tile HSX = B+2, out B^3).
// ---------------------------------------------------------------------------
void orc_kernel_laplacian(int B, const float *tile, float dt, int clamp01, float *out) {
    const int HSX = B + 2, dy = HSX, dz = HSX * HSX;
    i64 o = 0;
    for (int z = 1; z <= B; ++z)
        for (int y = 1; y <= B; ++y)
            for (int x = 1; x <= B; ++x) {
                int h = z * dz + y * dy + x;
                float c = tile[h], lx = tile[h - 1], rx = tile[h + 1];
                float dyu = tile[h - dy], dyd = tile[h + dy], dzb = tile[h - dz], dzf = tile[h + dz];
                float nb = lx + rx + dyu + dyd + dzb + dzf;
                float lap = fmaf(c, -6.0f, nb);
                float fx = (rx - lx) * 0.5f, fy = (dyd - dyu) * 0.5f, fz = (dzf - dzb) * 0.5f;
                float gm = sqrtf(fmaf(fx, fx, fmaf(fy, fy, fz * fz)));
                float v = fmaf(dt * lap, gm, c);
                if (clamp01) v = clampf(v, 0.0f, 1.0f);
                out[o++] = v;
            }
}

// ---------------------------------------------------------------------------
// A9: kernel_meancurv -- math/meancurv.rs:17-107 (tile HSX = B+2).
// ---------------------------------------------------------------------------
void orc_kernel_meancurv(int B, const float *tile, float dt, int clamp01, float *out) {
    const int HSX = B + 2, dy = HSX, dz = HSX * HSX;
    i64 o = 0;
    for (int z = 1; z <= B; ++z)
        for (int y = 1; y <= B; ++y)
            for (int x = 1; x <= B; ++x) {
                int h = z * dz + y * dy + x;
                float f0 = tile[h], f1 = tile[h - 1], f2 = tile[h + 1], f3 = tile[h - dy], f4 = tile[h + dy];
                float f5 = tile[h - dz], f6 = tile[h + dz];
                float fx = (f2 - f1) * 0.5f, fy = (f4 - f3) * 0.5f, fz = (f6 - f5) * 0.5f;
                float fxx = f1 + f2 - 2.0f * f0, fyy = f3 + f4 - 2.0f * f0, fzz = f5 + f6 - 2.0f * f0;
                float d1 = tile[h + 1 + dy], d2 = tile[h + 1 - dy], d3 = tile[h - 1 + dy], d4 = tile[h - 1 - dy];
                float d5 = tile[h + 1 + dz], d6 = tile[h + 1 - dz], d7 = tile[h - 1 + dz], d8 = tile[h - 1 - dz];
                float d9 = tile[h + dy + dz], d10 = tile[h + dy - dz], d11 = tile[h - dy + dz], d12 = tile[h - dy - dz];
                float fyx = (d1 - d2 - d3 + d4) * 0.25f;
                float fzx = (d5 - d6 - d7 + d8) * 0.25f;
                float fzy = (d9 - d10 - d11 + d12) * 0.25f;
                float cross = fx * (fy * fyx) + fx * (fz * fzx) + fy * (fz * fzy);
                float hnum = fmaf(fy * fy + fz * fz, fxx,
                                  fmaf(fx * fx + fz * fz, fyy, fmaf(fx * fx + fy * fy, fzz, -2.0f * cross)));
                float grad = fmaf(fx, fx, fmaf(fy, fy, fz * fz));
                float hc = grad > 1e-7f ? hnum / grad : 0.0f;
                float d = clampf(dt * hc, -0.1f, 0.1f);
                float v = f0 + d;
                if (clamp01) v = clampf(v, 0.0f, 1.0f);
                out[o++] = v;
            }
}

// ---------------------------------------------------------------------------
// A10: kernel_bilaplacian -- math/bilaplacian.rs:15-97 (tile HSX = B+4).
// ---------------------------------------------------------------------------
void orc_kernel_bilaplacian(int B, const float *tile, float dt, int clamp01, float *out) {
    const int HSX = B + 4, dy = HSX, dz = HSX * HSX;
    i64 o = 0;
    for (int z = 2; z <= B + 1; ++z)
        for (int y = 2; y <= B + 1; ++y)
            for (int x = 2; x <= B + 1; ++x) {
                int h = z * dz + y * dy + x;
                const float *t = tile;
                float f0 = t[h], f1 = t[h - 1], f2 = t[h + 1], f3 = t[h - dy], f4 = t[h + dy], f5 = t[h - dz],
                      f6 = t[h + dz];
                float g1 = t[h - 2], g2 = t[h + 2], g3 = t[h - 2 * dy], g4 = t[h + 2 * dy], g5 = t[h - 2 * dz],
                      g6 = t[h + 2 * dz];
                float s1 = f1 + f2 + f3 + f4 + f5 + f6;
                float s2 = g1 + g2 + g3 + g4 + g5 + g6;
                float sd = t[h + 1 + dy] + t[h + 1 - dy] + t[h - 1 + dy] + t[h - 1 - dy] + t[h + 1 + dz] +
                           t[h + 1 - dz] + t[h - 1 + dz] + t[h - 1 - dz] + t[h + dy + dz] + t[h + dy - dz] +
                           t[h - dy + dz] + t[h - dy - dz];
                float hh = 42.0f * f0 - 12.0f * s1 + s2 + 2.0f * sd;
                float fx = (f2 - f1) * 0.5f, fy = (f4 - f3) * 0.5f, fz = (f6 - f5) * 0.5f;
                float gm = sqrtf(fmaf(fx, fx, fmaf(fy, fy, fz * fz)));
                float upd = (dt * hh) * gm;
                float d = clampf(-upd, -0.1f, 0.1f);
                float v = f0 + d;
                if (clamp01) v = clampf(v, 0.0f, 1.0f);
                out[o++] = v;
            }
}

// ---------------------------------------------------------------------------
// A5: Sweeper::sweep / run -- solver.rs:166-217, over f32 storage with the
// Neumann boundary (solver.rs:205-209).  Colour passes run in order; blocks
// inside one colour are independent (solver.rs:242-246) and run on `nthreads`
// OpenMP threads, each with its own halo tile.  Dummy/absent blocks are
// skipped (solver.rs:201-204).  stencil: 0 Laplacian, 1 MeanCurvature,
// 2 BiLaplacian.  `ftz` turns on FTZ/DAZ as thread_pool.rs:5-11 does.
// ---------------------------------------------------------------------------
void orc_sweep(i64 sx, i64 sy, i64 sz, float *slab, const i32 *b2s, const i64 *active, i64 n, int stencil,
               float dt, int iterations, int clamp01, int nthreads, int ftz) {
    const int L = 4, B = 16;
    std::vector<i64> order(n > 0 ? n : 1);
    i64 counts[8];
    orc_color_buckets(sx, sy, sz, L, active, n, order.data(), counts);
    const int halo = stencil == 2 ? 2 : 1;
    const int full = stencil == 0 ? 0 : 1;
    const int HSX = B + 2 * halo;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#endif
    for (int it = 0; it < iterations; ++it) {
        i64 off = 0;
        for (int c = 0; c < 8; ++c) {
            const i64 *bucket = order.data() + off;
            const i64 cnt = counts[c];
#pragma omp parallel num_threads(nthreads)
            {
                if (ftz) {
                    _MM_SET_FLUSH_ZERO_MODE(_MM_FLUSH_ZERO_ON);
                    _MM_SET_DENORMALS_ZERO_MODE(_MM_DENORMALS_ZERO_ON);
                }
                std::vector<float> tile((size_t)HSX * HSX * HSX);
#pragma omp for schedule(dynamic, 4)
                for (i64 k = 0; k < cnt; ++k) {
                    i64 bid = bucket[k];
                    i32 s = b2s[bid];
                    if (s < 2) continue;
                    float *out = slab + (i64)s * 4096;
                    orc_halo_fill(sx, sy, sz, L, slab, b2s, bid, halo, full, BC_NEUMANN, tile.data());
                    if (stencil == 0)
                        orc_kernel_laplacian(B, tile.data(), dt, clamp01, out);
                    else if (stencil == 1)
                        orc_kernel_meancurv(B, tile.data(), dt, clamp01, out);
                    else
                        orc_kernel_bilaplacian(B, tile.data(), dt, clamp01, out);
                }
                if (ftz) {
                    _MM_SET_FLUSH_ZERO_MODE(_MM_FLUSH_ZERO_OFF);
                    _MM_SET_DENORMALS_ZERO_MODE(_MM_DENORMALS_ZERO_OFF);
                }
            }
            off += cnt;
        }
    }
}

// Sweeper::sweep for D = Density (u16): the halo is filled from the
// dequantised field u * (1/65535) (Density::to_f32, channel.rs:124-128;
// halo.rs:95) and every smoothed value is stored as round(v * 65535), half
// away from zero (StoreBack<Density>, math/stencil.rs:162-174), saturated to
// [0, 65535].  A FP32 mirror of the dequantised slab carries the in-place
// updates between colour passes.
void orc_sweep_density(i64 sx, i64 sy, i64 sz, uint16_t *slab16, i64 n_slots, const i32 *b2s, const i64 *active,
                       i64 n, int stencil, float dt, int iterations, int clamp01) {
    const int L = 4, B = 16;
    const float inv = 1.0f / 65535.0f;
    std::vector<float> f((size_t)n_slots * 4096);
    for (size_t i = 0; i < f.size(); ++i) f[i] = (float)slab16[i] * inv;
    std::vector<i64> order(n > 0 ? n : 1);
    i64 counts[8];
    orc_color_buckets(sx, sy, sz, L, active, n, order.data(), counts);
    const int halo = stencil == 2 ? 2 : 1;
    const int full = stencil == 0 ? 0 : 1;
    const int HSX = B + 2 * halo;
    std::vector<float> tile((size_t)HSX * HSX * HSX), out(4096);
    for (int it = 0; it < iterations; ++it) {
        i64 off = 0;
        for (int c = 0; c < 8; ++c) {
            for (i64 k = 0; k < counts[c]; ++k) {
                const i64 bid = order[off + k];
                const i32 s = b2s[bid];
                if (s < 2) continue;
                orc_halo_fill(sx, sy, sz, L, f.data(), b2s, bid, halo, full, BC_NEUMANN, tile.data());
                std::memcpy(out.data(), &f[(size_t)s * 4096], 4096 * sizeof(float));
                if (stencil == 0)
                    orc_kernel_laplacian(B, tile.data(), dt, clamp01, out.data());
                else if (stencil == 1)
                    orc_kernel_meancurv(B, tile.data(), dt, clamp01, out.data());
                else
                    orc_kernel_bilaplacian(B, tile.data(), dt, clamp01, out.data());
                // cells past a partial domain end are written by the kernels too; store them all
                for (int v = 0; v < 4096; ++v) {
                    const float r = std::min(std::max(roundf(out[v] * 65535.0f), 0.0f), 65535.0f);
                    const uint16_t q = (uint16_t)(int)r;
                    slab16[(size_t)s * 4096 + v] = q;
                    f[(size_t)s * 4096 + v] = (float)q * inv;
                }
            }
            off += counts[c];
        }
    }
}

// Sweeper::sweep for D = Density8 (u8, channel.rs:55-59,155-189): the halo
// is filled from the linear dequantisation u * (1/255) (the Density8 Dequant,
// math/sample.rs:64-69) and every smoothed value v is stored sqrt-compressed
// with stochastic rounding (StoreBack<Density8>, math/stencil.rs:176-192):
// vf = sqrt(v) * 255, code = floor(vf) + (u < vf - floor(vf)), saturated to
// [0, 255].  The reference's draw u comes from per-thread SimdRng lanes
// (math/stencil.rs:101-132), a stream that depends on the SIMD width and the
// schedule; this restatement uses the stream the C ABI states for it, a fixed
// function of (seed, sweep, bid, voxel) built from SimdRng's own pieces (the
// murmur3 finaliser of SimdRng::seed and the x16807 step of SimdRng::next).
// Sweep k of the call is sweep index sweep0 + k.
static inline u32 d8_fmix(u32 h) {
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

float orc_density8_uniform(u32 seed, u32 sweep, u32 bid, u32 vid) {
    const u32 key = d8_fmix(seed ^ (sweep * 0x9e3779b9u));
    u32 h = d8_fmix(d8_fmix(key ^ bid) ^ (vid * 0x9e3779b9u));
    if (h == 0u) h = 0xdeadbeefu;
    const u32 bits = ((h * 16807u) >> 9) | 0x3f800000u;
    float f;
    std::memcpy(&f, &bits, 4);
    return f - 1.0f;
}

uint8_t orc_density8_code(float v, float u) {
    const float vf = sqrtf(v) * 255.0f;
    const float fl = floorf(vf);
    const float r = u < vf - fl ? fl + 1.0f : fl;
    if (r >= 255.0f) return 255;
    return r > 0.0f ? (uint8_t)(int)r : 0;
}

void orc_sweep_density8(i64 sx, i64 sy, i64 sz, uint8_t *slab8, i64 n_slots, const i32 *b2s, const i64 *active,
                        i64 n, int stencil, float dt, int iterations, int clamp01, u32 seed, u32 sweep0) {
    const int L = 4, B = 16;
    const float inv = 1.0f / 255.0f;
    std::vector<float> f((size_t)n_slots * 4096);
    for (size_t i = 0; i < f.size(); ++i) f[i] = (float)slab8[i] * inv;
    std::vector<i64> order(n > 0 ? n : 1);
    i64 counts[8];
    orc_color_buckets(sx, sy, sz, L, active, n, order.data(), counts);
    const int halo = stencil == 2 ? 2 : 1;
    const int full = stencil == 0 ? 0 : 1;
    const int HSX = B + 2 * halo;
    std::vector<float> tile((size_t)HSX * HSX * HSX), out(4096);
    for (int it = 0; it < iterations; ++it) {
        i64 off = 0;
        for (int c = 0; c < 8; ++c) {
            for (i64 k = 0; k < counts[c]; ++k) {
                const i64 bid = order[off + k];
                const i32 s = b2s[bid];
                if (s < 2) continue;
                orc_halo_fill(sx, sy, sz, L, f.data(), b2s, bid, halo, full, BC_NEUMANN, tile.data());
                if (stencil == 0)
                    orc_kernel_laplacian(B, tile.data(), dt, clamp01, out.data());
                else if (stencil == 1)
                    orc_kernel_meancurv(B, tile.data(), dt, clamp01, out.data());
                else
                    orc_kernel_bilaplacian(B, tile.data(), dt, clamp01, out.data());
                for (int v = 0; v < 4096; ++v) {
                    const float u = orc_density8_uniform(seed, sweep0 + (u32)it, (u32)bid, (u32)v);
                    const uint8_t q = orc_density8_code(out[v], u);
                    slab8[(size_t)s * 4096 + v] = q;
                    f[(size_t)s * 4096 + v] = (float)q * inv;
                }
            }
            off += counts[c];
        }
    }
}

// Multires sweep: the same colour schedule over the level-l blocks in
// `active`, with the halo gathered by fill_multires (halo.rs:187) from level l
// and level l+1.  The reference runs fill_multires only from benches
// (multires_bench.rs:159-169); this is the sweep those benches time, made a
// call.
// FTZ / DAZ inside orc_sweep_multires' worker threads (the benches' pool
// setting, thread_pool.rs:5-11); off by default, as the reference's tests run.
static int g_multires_ftz = 0;
void orc_set_multires_ftz(int on) { g_multires_ftz = on; }

void orc_sweep_multires(i64 sx, i64 sy, i64 sz, int L, float *fine_slab, const i32 *fine_b2s,
                        const float *coarse_slab, const i32 *coarse_b2s, const uint8_t *levels, const i64 *active,
                        i64 n, int stencil, float dt, int iterations, int clamp01, int nthreads) {
    const int B = 1 << L;
    const i64 N = (i64)B * B * B;
    std::vector<i64> order(n > 0 ? n : 1);
    i64 counts[8];
    orc_color_buckets(sx, sy, sz, L, active, n, order.data(), counts);
    const int halo = stencil == 2 ? 2 : 1;
    const int full = stencil == 0 ? 0 : 1;
    const int HSX = B + 2 * halo;
    for (int it = 0; it < iterations; ++it) {
        i64 off = 0;
        for (int c = 0; c < 8; ++c) {
            const i64 *bucket = order.data() + off;
            const i64 cnt = counts[c];
#pragma omp parallel num_threads(nthreads < 1 ? 1 : nthreads)
            {
                const int ftz = g_multires_ftz;
                if (ftz) {
                    _MM_SET_FLUSH_ZERO_MODE(_MM_FLUSH_ZERO_ON);
                    _MM_SET_DENORMALS_ZERO_MODE(_MM_DENORMALS_ZERO_ON);
                }
                std::vector<float> tile((size_t)HSX * HSX * HSX);
#pragma omp for schedule(dynamic, 4)
                for (i64 k = 0; k < cnt; ++k) {
                    i64 bid = bucket[k];
                    i32 s = fine_b2s[bid];
                    if (s < 2) continue;
                    orc_halo_fill_multires(sx, sy, sz, L, fine_slab, fine_b2s, coarse_slab, coarse_b2s, levels, bid,
                                           halo, full, BC_NEUMANN, tile.data());
                    float *out = fine_slab + (i64)s * N;
                    if (stencil == 0)
                        orc_kernel_laplacian(B, tile.data(), dt, clamp01, out);
                    else if (stencil == 1)
                        orc_kernel_meancurv(B, tile.data(), dt, clamp01, out);
                    else
                        orc_kernel_bilaplacian(B, tile.data(), dt, clamp01, out);
                }
                if (ftz) {
                    _MM_SET_FLUSH_ZERO_MODE(_MM_FLUSH_ZERO_OFF);
                    _MM_SET_DENORMALS_ZERO_MODE(_MM_DENORMALS_ZERO_OFF);
                }
            }
            off += cnt;
        }
    }
}

// ---------------------------------------------------------------------------
// Halos of blocks at coarse levels (no reference counterpart: fill_multires,
// halo.rs:187-358, only fills a fine centre block).  The centre block `bid`
// sits at level l >= 1 (block edge E = 2^L[l]) of a multires grid whose level
// k has dims[3k..3k+2] voxels, slab slabs[k] and slot table b2s[k] (block edge
// halving per level, one block grid).  Tile cell (hx, hy, hz) of the HSX^3
// tile (HSX = E + 2 halo) holds level-l voxel g = block origin + h - halo,
// each axis resolved by the Neumann mirror of that level (boundary.rs:48-56);
// the voxel's block b then decides where it is read, by levels[b]:
//   * levels[b] == l: level l's own voxel;
//   * levels[b] >  l: level l+1's voxel g >> 1 (the injection fill_multires
//     uses for a coarser neighbour, halo.rs:263-268);
//   * levels[b] <  l: the mean of level l-1's 2^3 voxels 2g + (dx, dy, dz),
//     summed dz, dy, dx in order and divided by 8 (downsample_channel_avg,
//     downsample.rs:44-53).
// A block absent at the level read gives that level's empty value (slot 0).
// ---------------------------------------------------------------------------
static float level_read(const i64 *dims, const float *const *slabs, const i32 *const *b2s, int k, i32 x, i32 y,
                        i32 z) {
    const int L = 4 - k;
    Level g(dims[3 * k], dims[3 * k + 1], dims[3 * k + 2], L, slabs[k], b2s[k]);
    const i32 m = g.B - 1;
    const i64 bid = (x >> L) + (i64)(y >> L) * g.nx + (i64)(z >> L) * g.nxy;
    const i64 vid = (x & m) | ((i64)(y & m) << L) | ((i64)(z & m) << (2 * L));
    return g.block_ptr(bid)[vid];
}

void orc_halo_fill_level(const i64 *dims, const float *const *slabs, const i32 *const *b2s, const uint8_t *levels,
                         int n_levels, int l, i64 bid, int halo, float *tile) {
    const int L = 4 - l, E = 1 << L, HSX = E + 2 * halo;
    Level g(dims[3 * l], dims[3 * l + 1], dims[3 * l + 2], L, slabs[l], b2s[l]);
    const i32 bx0 = (i32)(bid % g.nx), by0 = (i32)((bid / g.nx) % g.ny), bz0 = (i32)(bid / g.nxy);
    for (int hz = 0; hz < HSX; ++hz)
        for (int hy = 0; hy < HSX; ++hy)
            for (int hx = 0; hx < HSX; ++hx) {
                i32 rx, ry, rz;
                resolve_axis(bx0 * E + hx - halo, (i32)g.sx - 1, BC_NEUMANN, &rx);
                resolve_axis(by0 * E + hy - halo, (i32)g.sy - 1, BC_NEUMANN, &ry);
                resolve_axis(bz0 * E + hz - halo, (i32)g.sz - 1, BC_NEUMANN, &rz);
                const i64 b = (rx >> L) + (i64)(ry >> L) * g.nx + (i64)(rz >> L) * g.nxy;
                float v;
                if (levels[b] == l || (levels[b] > l && l + 1 >= n_levels) || (levels[b] < l && l == 0)) {
                    v = level_read(dims, slabs, b2s, l, rx, ry, rz);
                } else if (levels[b] > l) {
                    v = level_read(dims, slabs, b2s, l + 1, rx >> 1, ry >> 1, rz >> 1);
                } else {
                    float sum = 0.0f;
                    for (int dz = 0; dz < 2; ++dz)
                        for (int dy = 0; dy < 2; ++dy)
                            for (int dx = 0; dx < 2; ++dx)
                                sum += level_read(dims, slabs, b2s, l - 1, 2 * rx + dx, 2 * ry + dy, 2 * rz + dz);
                    v = sum / 8.0f;
                }
                tile[((size_t)hz * HSX + hy) * HSX + hx] = v;
            }
}

// The 8-colour sweep of a multires grid's blocks at all their levels: the
// colour of a block is the parity of its block coordinates whatever its level
// (solver.rs:247-267), and a colour pass updates every active block of that
// colour -- at its own level, in place -- before the next colour starts.  A
// level-0 block fills its halo with fill_multires (orc_halo_fill_multires,
// levels 0 and 1); a block at level l >= 1 with orc_halo_fill_level.  Blocks
// of one colour are never neighbours and a halo reaches only the 26
// neighbours (halo <= E), so the order inside a colour does not matter.  The
// kernels are the reference's (laplacian.rs, meancurv.rs, bilaplacian.rs) on
// the level's block edge.  Blocks without a value slot at their level are
// skipped (solver.rs:201-204).
void orc_sweep_levels(const i64 *dims, float *const *slabs, const i32 *const *b2s, const uint8_t *levels,
                      int n_levels, const i64 *active, i64 n, int stencil, float dt, int iterations, int clamp01,
                      int nthreads) {
    const int halo = stencil == 2 ? 2 : 1;
    const int full = stencil == 0 ? 0 : 1;
    Level g0(dims[0], dims[1], dims[2], 4, nullptr, nullptr);
    std::vector<i64> by_color[8];
    for (i64 i = 0; i < n; ++i) {
        const i64 bid = active[i];
        const i64 bx = bid % g0.nx, by = (bid / g0.nx) % g0.ny, bz = bid / g0.nxy;
        by_color[(bx & 1) | ((by & 1) << 1) | ((bz & 1) << 2)].push_back(bid);
    }
    for (int it = 0; it < iterations; ++it)
        for (int c = 0; c < 8; ++c) {
            const std::vector<i64> &bucket = by_color[c];
#pragma omp parallel num_threads(nthreads < 1 ? 1 : nthreads)
            {
                std::vector<float> tile(20 * 20 * 20);
#pragma omp for schedule(dynamic, 16)
                for (i64 k = 0; k < (i64)bucket.size(); ++k) {
                    const i64 bid = bucket[k];
                    const int l = levels[bid];
                    const i32 s = b2s[l][bid];
                    if (s < 2) continue;
                    const int L = 4 - l, E = 1 << L;
                    float *out = slabs[l] + (i64)s * E * E * E;
                    if (l == 0)
                        orc_halo_fill_multires(dims[0], dims[1], dims[2], 4, slabs[0], b2s[0], slabs[1], b2s[1],
                                               levels, bid, halo, full, BC_NEUMANN, tile.data());
                    else
                        orc_halo_fill_level(dims, slabs, b2s, levels, n_levels, l, bid, halo, tile.data());
                    if (stencil == 0)
                        orc_kernel_laplacian(E, tile.data(), dt, clamp01, out);
                    else if (stencil == 1)
                        orc_kernel_meancurv(E, tile.data(), dt, clamp01, out);
                    else
                        orc_kernel_bilaplacian(E, tile.data(), dt, clamp01, out);
                }
            }
        }
}

// ---------------------------------------------------------------------------
// A16: downsample_channel_avg -- multires/downsample.rs:13-58.  For every bid
// in `bids`, coarse[b] = mean of the 8 children, summed dz, dy, dx in order
// and divided by 8.0 (downsample.rs:44-53).  coarse_b2s[bid] must be >= 2.
// ---------------------------------------------------------------------------
void orc_downsample_avg(int L, const float *fine_slab, const i32 *fine_b2s, float *coarse_slab,
                        const i32 *coarse_b2s, const i64 *bids, i64 n) {
    const int Bc = 1 << (L - 1);
    const i64 Nf = (i64)1 << (3 * L), Nc = (i64)Bc * Bc * Bc;
    for (i64 k = 0; k < n; ++k) {
        const float *sp = fine_slab + (i64)fine_b2s[bids[k]] * Nf;
        float *dp = coarse_slab + (i64)coarse_b2s[bids[k]] * Nc;
        i64 vid = 0;
        for (int cz = 0; cz < Bc; ++cz)
            for (int cy = 0; cy < Bc; ++cy)
                for (int cx = 0; cx < Bc; ++cx) {
                    float sum = 0.0f;
                    for (int dz = 0; dz < 2; ++dz)
                        for (int dy = 0; dy < 2; ++dy)
                            for (int dx = 0; dx < 2; ++dx) {
                                i64 i = (2 * cx + dx) | ((i64)(2 * cy + dy) << L) | ((i64)(2 * cz + dz) << (2 * L));
                                sum += sp[i];
                            }
                    dp[vid++] = sum / 8.0f;
                }
    }
}

// downsample.rs:67-82 -- coarse ghost sample at a fine coordinate.
float orc_sample_coarse(i64 csx, i64 csy, i64 csz, int Lc, const float *cslab, const i32 *cb2s, i32 gx, i32 gy,
                        i32 gz) {
    Level c(csx, csy, csz, Lc, cslab, cb2s);
    i32 x = std::min(std::max(gx >> 1, 0), (i32)csx - 1);
    i32 y = std::min(std::max(gy >> 1, 0), (i32)csy - 1);
    i32 z = std::min(std::max(gz >> 1, 0), (i32)csz - 1);
    i32 m = c.B - 1;
    i64 bid = (x >> Lc) + (i64)(y >> Lc) * c.nx + (i64)(z >> Lc) * c.nxy;
    i64 vid = (x & m) | ((i64)(y & m) << Lc) | ((i64)(z & m) << (2 * Lc));
    return c.block_ptr(bid)[vid];
}

// ---------------------------------------------------------------------------
// regularize_refinement_map -- multires/refinement.rs:66-107 (2:1 balance).
// ---------------------------------------------------------------------------
void orc_regularize(uint8_t *map, i64 nx, i64 ny, i64 nz, int n_levels) {
    if (n_levels <= 2) return;
    const i64 nxy = nx * ny, nb = nxy * nz;
    std::vector<uint8_t> prev(map, map + nb);
    for (int ll = 1; ll <= n_levels - 2; ++ll) {
        for (i64 bid = 0; bid < nb; ++bid) {
            uint8_t level = prev[bid], out = level;
            if (level >= ll) {
                i64 bx = bid % nx, by = (bid / nx) % ny, bz = bid / nxy;
                uint8_t level2 = (uint8_t)(n_levels - 1);
                for (int dz = -1; dz <= 1; ++dz)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dx = -1; dx <= 1; ++dx) {
                            if (!dx && !dy && !dz) continue;
                            i64 x2 = bx + dx, y2 = by + dy, z2 = bz + dz;
                            if (x2 < 0 || x2 >= nx || y2 < 0 || y2 >= ny || z2 < 0 || z2 >= nz) continue;
                            level2 = std::min(level2, prev[x2 + y2 * nx + z2 * nxy]);
                        }
                if (level2 == ll - 1) out = (uint8_t)ll;
            }
            map[bid] = out;
        }
        std::memcpy(prev.data(), map, nb);
    }
}

// ---------------------------------------------------------------------------
// A14: active_blocks -- particles/active.rs:15-47 with footprint_axis
// (particles/mod.rs:207-212): trunc(p/bsx -+ r_scan/bsx) clipped to [0,n).
// Output sorted ascending, deduplicated; returns the count.
// ---------------------------------------------------------------------------
i64 orc_active_blocks(i64 sx, i64 sy, i64 sz, const float *pos, i64 n, float r_scan, i64 *out) {
    const i64 bsx = 16;
    const i64 nx = (sx + 15) >> 4, ny = (sy + 15) >> 4, nz = (sz + 15) >> 4, nxy = nx * ny;
    std::vector<uint8_t> flag((size_t)(nxy * nz), 0);
    const float rb = r_scan / (float)bsx;
    auto fp = [&](float p, i64 nn, i32 *lo, i32 *hi) {
        float b = p / (float)bsx;
        *lo = std::max((i32)truncf(b - rb), 0);
        *hi = std::min((i32)truncf(b + rb), (i32)nn - 1);
    };
    for (i64 i = 0; i < n; ++i) {
        i32 x1, x2, y1, y2, z1, z2;
        fp(pos[3 * i], nx, &x1, &x2);
        fp(pos[3 * i + 1], ny, &y1, &y2);
        fp(pos[3 * i + 2], nz, &z1, &z2);
        for (i32 bz = z1; bz <= z2; ++bz)
            for (i32 by = y1; by <= y2; ++by)
                for (i32 bx = x1; bx <= x2; ++bx) flag[bx + by * nx + bz * nxy] = 1;
    }
    i64 k = 0;
    for (i64 b = 0; b < nxy * nz; ++b)
        if (flag[b]) out[k++] = b;
    return k;
}

// ---------------------------------------------------------------------------
// A2: bucket_indices -- particles/sort.rs:174-272, in the stable form of the
// reference's own dense oracle bucket_by_block_dense (particles/sort.rs:455-490):
// perm is stable by block id, ties by original index.  Returns the number of
// occupied blocks; particle_blocks (ascending) and starts (len nblk+1).
// ---------------------------------------------------------------------------
i64 orc_bucket_indices(const i64 *bids, i64 n, i64 n_blocks, i64 *particle_blocks, i64 *starts, i64 *perm) {
    std::vector<i64> counts((size_t)n_blocks, 0);
    for (i64 j = 0; j < n; ++j) counts[bids[j]]++;
    std::vector<i64> block_start((size_t)n_blocks + 1);
    i64 acc = 0;
    for (i64 b = 0; b < n_blocks; ++b) {
        block_start[b] = acc;
        acc += counts[b];
    }
    block_start[n_blocks] = acc;
    std::vector<i64> cursor(block_start.begin(), block_start.end() - 1);
    for (i64 j = 0; j < n; ++j) perm[cursor[bids[j]]++] = j;
    i64 k = 0;
    for (i64 b = 0; b < n_blocks; ++b)
        if (counts[b] > 0) {
            particle_blocks[k] = b;
            starts[k] = block_start[b];
            ++k;
        }
    starts[k] = block_start[n_blocks];
    return k;
}

// Rust `f32 as usize` (p2g.rs:153): truncation, saturating at 0 below zero
// and for NaN (a plain C++ cast of a negative float to unsigned is undefined).
static inline u64 as_usize(float p) { return p > 0.0f ? (u64)p : 0; }

// Row-major block id of a particle: p2g.rs:149-157.
void orc_particle_bids(i64 sx, i64 sy, i64 sz, const float *pos, i64 n, i64 *bids) {
    const i64 nx = (sx + 15) >> 4, ny = (sy + 15) >> 4;
    for (i64 i = 0; i < n; ++i) {
        u64 x = as_usize(pos[3 * i]), y = as_usize(pos[3 * i + 1]), z = as_usize(pos[3 * i + 2]);
        bids[i] = (i64)((x >> 4) + (y >> 4) * nx + (z >> 4) * nx * ny);
    }
}

// ---------------------------------------------------------------------------
// Particle Morton keys: morton3 (multires/sort.rs:20) of the particle's voxel
// (truncated position, p2g.rs:153), kept to its low 30 bits -- the full code
// for voxel coordinates < 1024, which is the 30-bit key the particle sort uses.
// ---------------------------------------------------------------------------
void orc_morton_keys(const float *pos, i64 n, u32 *keys) {
    for (i64 i = 0; i < n; ++i) {
        u64 k = orc_morton3(as_usize(pos[3 * i]), as_usize(pos[3 * i + 1]), as_usize(pos[3 * i + 2]));
        keys[i] = (u32)(k & 0x3fffffffULL);
    }
}

// Stable sort of indices by 32-bit key (std::stable_sort): the permutation a
// stable LSD radix sort must reproduce.
void orc_stable_sort_perm(const u32 *keys, i64 n, u32 *perm) {
    std::vector<u32> idx((size_t)n);
    std::iota(idx.begin(), idx.end(), 0u);
    std::stable_sort(idx.begin(), idx.end(), [&](u32 a, u32 b) { return keys[a] < keys[b]; });
    std::memcpy(perm, idx.data(), sizeof(u32) * (size_t)n);
}

}  // extern "C"

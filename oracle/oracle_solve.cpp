// CPU oracle, part 3: north-star methods with no reference counterpart --
// B-spline particle transfers and the Poisson relaxation / residual.
//
// TEST INFRASTRUCTURE ONLY (see oracle.cpp).  These functions define the
// semantics the HIP kernels implement; they are pinned in tests by closed
// forms (partition of unity, conservation, affine reproduction, brute-force
// scatter, fixed points), not by a reference implementation.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <unordered_map>
#include <vector>

typedef int64_t i64;
typedef int32_t i32;
typedef uint32_t u32;
typedef uint64_t u64;

extern "C" u64 orc_morton3(u64 x, u64 y, u64 z);

namespace {

inline u64 as_usize(float p) { return p > 0.0f ? (u64)p : 0; }

// Per-axis B-spline weight of the cell with index ci (centre ci + 1/2) for a
// particle at coordinate p.  order 1: tent max(0, 1-|x|); order 2: quadratic
// B-spline 3/4 - x^2 (|x| < 1/2), (3/2 - |x|)^2 / 2 (|x| < 3/2); x = p-1/2-ci.
inline float bspline_w(int order, float p, i64 ci) {
    const float x = (p - 0.5f) - (float)ci;
    const float ax = fabsf(x);
    if (order == 1) return ax < 1.0f ? 1.0f - ax : 0.0f;
    if (ax < 0.5f) return 0.75f - x * x;
    if (ax < 1.5f) {
        const float t = 1.5f - ax;
        return 0.5f * t * t;
    }
    return 0.0f;
}

}  // namespace

extern "C" {

// Per-axis weight, exposed for the closed-form tests.
float orc_bspline_weight(int order, float p, i64 ci) { return bspline_w(order, p, ci); }

// ---------------------------------------------------------------------------
// B-spline P2G (order 1 trilinear, order 2 quadratic), cell-centred grid of
// 16^3 blocks.  Particles are stably sorted by the 30-bit Morton key of their
// voxel; every cell of the `active` blocks then sums, over its 27 neighbour
// voxels (z, y, x from -1 to 1) and the particles of each voxel in sorted
// order, wm = (wx*wy)*wz * m into mass and wm*u_a into momentum a.  Output
// slabs (4 x n_slots x 4096 over b2s): mass, then velocity x/y/z = momentum /
// mass (0 where mass is 0).  mass may be null (m = 1).
// ---------------------------------------------------------------------------
void orc_p2g_bspline(int order, i64 sx, i64 sy, i64 sz, const float *pos, const float *vel, const float *mass, i64 n,
                     const i32 *b2s, i64 n_slots, const i64 *active, i64 n_active, float *slabs) {
    const i64 nx = (sx + 15) >> 4, ny = (sy + 15) >> 4;
    std::vector<u32> keys((size_t)n), perm((size_t)n);
    for (i64 i = 0; i < n; ++i)
        keys[i] = (u32)(orc_morton3(as_usize(pos[3 * i]), as_usize(pos[3 * i + 1]), as_usize(pos[3 * i + 2])) &
                        0x3fffffffULL);
    std::iota(perm.begin(), perm.end(), 0u);
    std::stable_sort(perm.begin(), perm.end(), [&](u32 a, u32 b) { return keys[a] < keys[b]; });
    // voxel (x + y*sx + z*sx*sy) -> [begin, end) in sorted order
    std::unordered_map<i64, std::pair<i64, i64>> vox;
    for (i64 k = 0; k < n;) {
        i64 e = k;
        while (e < n && keys[perm[e]] == keys[perm[k]]) ++e;
        const u32 j = perm[k];
        const i64 v = (i64)as_usize(pos[3 * j]) + (i64)as_usize(pos[3 * j + 1]) * sx +
                      (i64)as_usize(pos[3 * j + 2]) * sx * sy;
        vox[v] = {k, e};
        k = e;
    }
    std::memset(slabs, 0, sizeof(float) * 4 * n_slots * 4096);
    for (i64 a = 0; a < n_active; ++a) {
        const i64 bid = active[a];
        const i32 s = b2s[bid];
        if (s < 2) continue;
        const i64 bx = bid % nx, by = (bid / nx) % ny, bz = bid / (nx * ny);
        for (int vid = 0; vid < 4096; ++vid) {
            const i64 cx = bx * 16 + (vid & 15), cy = by * 16 + ((vid >> 4) & 15), cz = bz * 16 + (vid >> 8);
            float am = 0.0f, au[3] = {0.0f, 0.0f, 0.0f};
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const i64 vx = cx + dx, vy = cy + dy, vz = cz + dz;
                        if (vx < 0 || vy < 0 || vz < 0 || vx >= sx || vy >= sy || vz >= sz) continue;
                        auto it = vox.find(vx + vy * sx + vz * sx * sy);
                        if (it == vox.end()) continue;
                        for (i64 k = it->second.first; k < it->second.second; ++k) {
                            const u32 j = perm[k];
                            const float w = (bspline_w(order, pos[3 * j], cx) * bspline_w(order, pos[3 * j + 1], cy)) *
                                            bspline_w(order, pos[3 * j + 2], cz);
                            const float wm = w * (mass ? mass[j] : 1.0f);
                            am = am + wm;
                            au[0] = au[0] + wm * vel[3 * j];
                            au[1] = au[1] + wm * vel[3 * j + 1];
                            au[2] = au[2] + wm * vel[3 * j + 2];
                        }
                    }
            slabs[(i64)s * 4096 + vid] = am;
            for (int c = 0; c < 3; ++c) slabs[((i64)(1 + c) * n_slots + s) * 4096 + vid] = am > 0.0f ? au[c] / am : 0.0f;
        }
    }
}

// B-spline G2P: u_p = sum over the support cells floor(p)-1 .. floor(p)+1 per
// axis (z, y, x outer to inner) of (wx*wy)*wz * u_cell for each of the 3
// velocity slabs (3 x n_slots x 4096); cells outside the domain are skipped,
// cells of absent blocks read the empty slot.
void orc_g2p_bspline(int order, i64 sx, i64 sy, i64 sz, const float *vel_slabs, i64 n_slots, const i32 *b2s,
                     const float *pos, i64 n, float *out) {
    const i64 nx = (sx + 15) >> 4, ny = (sy + 15) >> 4;
    for (i64 i = 0; i < n; ++i) {
        const float *p = pos + 3 * i;
        const i64 fx = (i64)floorf(p[0]), fy = (i64)floorf(p[1]), fz = (i64)floorf(p[2]);
        float acc[3] = {0.0f, 0.0f, 0.0f};
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const i64 cx = fx + dx, cy = fy + dy, cz = fz + dz;
                    if (cx < 0 || cy < 0 || cz < 0 || cx >= sx || cy >= sy || cz >= sz) continue;
                    const float w = (bspline_w(order, p[0], cx) * bspline_w(order, p[1], cy)) * bspline_w(order, p[2], cz);
                    const i64 bid = (cx >> 4) + (cy >> 4) * nx + (cz >> 4) * nx * ny;
                    const i64 vid = (cx & 15) | ((cy & 15) << 4) | ((cz & 15) << 8);
                    const i64 s = b2s[bid];
                    for (int c = 0; c < 3; ++c) acc[c] = acc[c] + w * vel_slabs[((i64)c * n_slots + s) * 4096 + vid];
                }
        for (int c = 0; c < 3; ++c) out[3 * i + c] = acc[c];
    }
}

int orc_version() { return 3; }

}  // extern "C"

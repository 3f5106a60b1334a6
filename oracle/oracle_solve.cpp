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

// ===========================================================================
// Poisson relaxation on the sparse block grid.
//
// Unknowns are the cells of the value blocks in `active`.  The 7-point
// operator (A x)_c = 6 x_c - sum of the 6 neighbours reads neighbours with the
// halo semantics of the smoothers: Neumann (mirror) outside the domain, the
// empty / full dummy value for absent / full blocks.  With the empty value 0
// the absent blocks around a band are a homogeneous Dirichlet boundary.
//
// Storage: 0 = f32, 1 = f16 (IEEE binary16), 2 = bf16; values are converted
// to f32 on load and rounded to nearest-even on store, arithmetic is f32.
// Slabs are slot-major, 4096 elements per slot, the layout of the grid.
// ===========================================================================
namespace {

inline u32 fbits(float f) {
    u32 u;
    std::memcpy(&u, &f, 4);
    return u;
}
inline float bitsf(u32 u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// IEEE binary32 -> binary16, round to nearest even (incl. subnormals, overflow to inf).
uint16_t f32_to_f16(float f) {
    const u32 u = fbits(f);
    const u32 sign = (u >> 16) & 0x8000u;
    const u32 exp = (u >> 23) & 0xffu;
    u32 man = u & 0x7fffffu;
    if (exp == 0xffu) return (uint16_t)(sign | 0x7c00u | (man ? 0x200u | (man >> 13) : 0u));
    int e = (int)exp - 127 + 15;
    if (e >= 31) return (uint16_t)(sign | 0x7c00u);
    if (e <= 0) {  // subnormal half (or zero)
        if (e < -10) return (uint16_t)sign;
        man |= 0x800000u;
        const int shift = 14 - e;  // 24-bit mantissa -> 10 - (1 - e) bits
        const u32 half_man = man >> shift;
        const u32 rem = man & ((1u << shift) - 1u);
        const u32 halfway = 1u << (shift - 1);
        u32 r = half_man + ((rem > halfway || (rem == halfway && (half_man & 1u))) ? 1u : 0u);
        return (uint16_t)(sign | r);
    }
    u32 h = ((u32)e << 10) | (man >> 13);
    const u32 rem = man & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;  // may carry into the exponent (inf)
    return (uint16_t)(sign | h);
}

float f16_to_f32(uint16_t h) {
    const u32 sign = (u32)(h & 0x8000u) << 16;
    const u32 exp = (h >> 10) & 0x1fu;
    const u32 man = h & 0x3ffu;
    if (exp == 0) {
        const float v = (float)man * (1.0f / 16777216.0f);  // man * 2^-24, exact
        return sign ? -v : v;
    }
    if (exp == 31) return bitsf(sign | 0x7f800000u | (man << 13));
    return bitsf(sign | ((exp - 15 + 127) << 23) | (man << 13));
}

uint16_t f32_to_bf16(float f) {
    u32 u = fbits(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

float bf16_to_f32(uint16_t h) { return bitsf((u32)h << 16); }

struct Field {
    int storage;
    void *p;
    float get(i64 i) const {
        if (storage == 0) return ((const float *)p)[i];
        if (storage == 1) return f16_to_f32(((const uint16_t *)p)[i]);
        return bf16_to_f32(((const uint16_t *)p)[i]);
    }
    void set(i64 i, float v) const {
        if (storage == 0)
            ((float *)p)[i] = v;
        else if (storage == 1)
            ((uint16_t *)p)[i] = f32_to_f16(v);
        else
            ((uint16_t *)p)[i] = f32_to_bf16(v);
    }
};

struct PGrid {
    i64 sx, sy, sz, nx, ny, nxy;
    const i32 *b2s;
    PGrid(i64 a, i64 b, i64 c, const i32 *m) : sx(a), sy(b), sz(c), b2s(m) {
        nx = (sx + 15) >> 4;
        ny = (sy + 15) >> 4;
        nxy = nx * ny;
    }
    // slab index of global cell (x, y, z) with Neumann mirroring outside
    i64 index(i64 x, i64 y, i64 z) const {
        x = x < 0 ? -x - 1 : (x >= sx ? 2 * sx - 1 - x : x);
        y = y < 0 ? -y - 1 : (y >= sy ? 2 * sy - 1 - y : y);
        z = z < 0 ? -z - 1 : (z >= sz ? 2 * sz - 1 - z : z);
        const i64 bid = (x >> 4) + (y >> 4) * nx + (z >> 4) * nxy;
        return (i64)b2s[bid] * 4096 + ((x & 15) | ((y & 15) << 4) | ((z & 15) << 8));
    }
};

// t = (b + s) - 6 x with s summed x-, x+, y-, y+, z-, z+ (the residual b - A x).
inline float residual_at(const PGrid &g, const Field &x, const Field &b, i64 gx, i64 gy, i64 gz, float *xc_out) {
    const float xc = x.get(g.index(gx, gy, gz));
    const float s = x.get(g.index(gx - 1, gy, gz)) + x.get(g.index(gx + 1, gy, gz)) + x.get(g.index(gx, gy - 1, gz)) +
                    x.get(g.index(gx, gy + 1, gz)) + x.get(g.index(gx, gy, gz - 1)) + x.get(g.index(gx, gy, gz + 1));
    *xc_out = xc;
    return fmaf(-6.0f, xc, b.get(g.index(gx, gy, gz)) + s);
}

template <typename F>
void for_cells(const PGrid &g, const i64 *active, i64 n, F f) {
    for (i64 k = 0; k < n; ++k) {
        const i64 bid = active[k];
        const i32 s = g.b2s[bid];
        if (s < 2) continue;
        const i64 bx = bid % g.nx, by = (bid / g.nx) % g.ny, bz = bid / g.nxy;
        for (int v = 0; v < 4096; ++v) {
            const i64 gx = bx * 16 + (v & 15), gy = by * 16 + ((v >> 4) & 15), gz = bz * 16 + (v >> 8);
            if (gx >= g.sx || gy >= g.sy || gz >= g.sz) continue;  // cells past a partial domain end
            f((i64)s * 4096 + v, gx, gy, gz);
        }
    }
}

}  // namespace

extern "C" {

// Storage conversions, exposed for the closed-form tests.
uint16_t orc_f32_to_f16(float f) { return f32_to_f16(f); }
float orc_f16_to_f32(uint16_t h) { return f16_to_f32(h); }
uint16_t orc_f32_to_bf16(float f) { return f32_to_bf16(f); }
float orc_bf16_to_f32(uint16_t h) { return bf16_to_f32(h); }

// Residual r = b - A x (r may be null, stored f32) and its FP64 squared norm;
// also the squared norm of b.
void orc_poisson_residual(i64 sx, i64 sy, i64 sz, int storage, const void *x, const void *b, const i32 *b2s,
                          const i64 *active, i64 n, float *r, double *r2, double *b2) {
    PGrid g(sx, sy, sz, b2s);
    Field X{storage, (void *)x}, Bf{storage, (void *)b};
    double rr = 0.0, bb = 0.0;
    for_cells(g, active, n, [&](i64 i, i64 gx, i64 gy, i64 gz) {
        float xc;
        const float t = residual_at(g, X, Bf, gx, gy, gz, &xc);
        if (r) r[i] = t;
        rr += (double)t * (double)t;
        const double bv = Bf.get(i);
        bb += bv * bv;
    });
    *r2 = rr;
    *b2 = bb;
}

// Residual of a compensated iterate x + xlo (FP32 fields, no reference
// counterpart): b, -6 x_c (split exactly by an fma) and the six x neighbours
// (x-, x+, y-, y+, z-, z+) summed with Knuth's TwoSum, their rounding errors
// accumulated in FP32; the xlo part as fmaf(-6, xlo_c, s_lo) with s_lo summed
// in the same neighbour order; r = sum + errors, rounded once.
void orc_poisson_residual_pair(i64 sx, i64 sy, i64 sz, const float *x, const float *xlo, const float *b,
                               const i32 *b2s, const i64 *active, i64 n, float *r, double *r2, double *b2) {
    PGrid g(sx, sy, sz, b2s);
    auto two_sum = [](float a, float c, float &err) {
        const float s = a + c;
        const float bb = s - a;
        err = (a - (s - bb)) + (c - bb);
        return s;
    };
    double rr = 0.0, bb = 0.0;
    for_cells(g, active, n, [&](i64 i, i64 gx, i64 gy, i64 gz) {
        const i64 q[7] = {g.index(gx, gy, gz),     g.index(gx - 1, gy, gz), g.index(gx + 1, gy, gz),
                          g.index(gx, gy - 1, gz), g.index(gx, gy + 1, gz), g.index(gx, gy, gz - 1),
                          g.index(gx, gy, gz + 1)};
        const float bi = b[i];
        const float p6 = 6.0f * x[q[0]];
        const float err = fmaf(6.0f, x[q[0]], -p6);
        float e;
        float acc = two_sum(bi, -p6, e);
        e -= err;
        for (int k = 1; k < 7; ++k) {
            float t;
            acc = two_sum(acc, x[q[k]], t);
            e += t;
        }
        const float slo = ((((xlo[q[1]] + xlo[q[2]]) + xlo[q[3]]) + xlo[q[4]]) + xlo[q[5]]) + xlo[q[6]];
        e += fmaf(-6.0f, xlo[q[0]], slo);
        const float t = acc + e;
        if (r) r[i] = t;
        rr += (double)t * (double)t;
        bb += (double)bi * (double)bi;
    });
    *r2 = rr;
    *b2 = bb;
}

// Weighted Jacobi: x_out = x + (omega/6) (b + s - 6 x) for every active cell.
void orc_jacobi(i64 sx, i64 sy, i64 sz, int storage, const void *x_in, void *x_out, const void *b, const i32 *b2s,
                const i64 *active, i64 n, float omega) {
    PGrid g(sx, sy, sz, b2s);
    Field X{storage, (void *)x_in}, Y{storage, x_out}, Bf{storage, (void *)b};
    const float w6 = omega / 6.0f;
    for_cells(g, active, n, [&](i64 i, i64 gx, i64 gy, i64 gz) {
        float xc;
        const float t = residual_at(g, X, Bf, gx, gy, gz, &xc);
        Y.set(i, fmaf(w6, t, xc));
    });
}

// One red-black SOR half sweep in place: cells with (x+y+z) & 1 == color.
void orc_rbgs(i64 sx, i64 sy, i64 sz, int storage, void *x, const void *b, const i32 *b2s, const i64 *active, i64 n,
              int color, float omega) {
    PGrid g(sx, sy, sz, b2s);
    Field X{storage, x}, Bf{storage, (void *)b};
    const float w6 = omega / 6.0f;
    for_cells(g, active, n, [&](i64 i, i64 gx, i64 gy, i64 gz) {
        if (((gx + gy + gz) & 1) != color) return;
        float xc;
        const float t = residual_at(g, X, Bf, gx, gy, gz, &xc);
        X.set(i, fmaf(w6, t, xc));
    });
}

// Conversions of a whole slab (n_elems) between storages (RNE on narrowing).
void orc_convert(int src_storage, const void *src, int dst_storage, void *dst, i64 n_elems) {
    Field S{src_storage, (void *)src}, D{dst_storage, dst};
    for (i64 i = 0; i < n_elems; ++i) D.set(i, S.get(i));
}

// The solve loop.  refine = 0: `per` red-black sweeps (colour 0 then 1) on x,
// then the residual check.  refine = 1: FP32 iterative refinement: from the
// last residual r (FP32) and its squared norm rr, 2^k = the power of two with
// rms(r) / 2^k in [1, 2) (k = 0 when r = 0), rt = T(r * 2^-k), e = 0, `per`
// inner red-black sweeps on A e = rt in storage T, x = fmaf(2^k, e, x), then
// the residual.  Stops when ||r|| <= tol ||b|| or when another period would
// exceed max_sweeps.  x (FP32 slab) is updated in place; out = {sweeps,
// checks, converged, rel_residual}.
void orc_poisson_solve(i64 sx, i64 sy, i64 sz, float *x, const float *b, i64 n_slots, const i32 *b2s,
                       const i64 *active, i64 n, int storage, int refine, float omega, int per, int max_sweeps,
                       double tol, double *out) {
    PGrid g(sx, sy, sz, b2s);
    const i64 N = n_slots * 4096;
    std::vector<float> r(N, 0.0f);
    const int eb = storage == 0 ? 4 : 2;
    std::vector<unsigned char> e((size_t)N * eb, 0), rt((size_t)N * eb, 0);
    i64 n_cells = 0;
    for_cells(g, active, n, [&](i64, i64, i64, i64) { ++n_cells; });
    double r2 = 0.0, b2 = 0.0;
    int sweeps = 0, checks = 1, converged = 0;
    orc_poisson_residual(sx, sy, sz, 0, x, b, b2s, active, n, r.data(), &r2, &b2);
    double rel = 0.0;
    for (;;) {
        rel = b2 > 0.0 ? std::sqrt(r2 / b2) : std::sqrt(r2);
        if (rel <= tol) {
            converged = 1;
            break;
        }
        if (!std::isfinite(rel) || sweeps + per > max_sweeps) break;
        if (!refine) {
            for (int k = 0; k < per; ++k) {
                orc_rbgs(sx, sy, sz, 0, x, b, b2s, active, n, 0, omega);
                orc_rbgs(sx, sy, sz, 0, x, b, b2s, active, n, 1, omega);
            }
        } else {
            int ex = 0;
            const double rms = n_cells > 0 ? std::sqrt(r2 / (double)n_cells) : 0.0;
            if (rms > 0.0 && std::isfinite(rms)) {
                std::frexp(rms, &ex);
                ex -= 1;
                ex = ex < -120 ? -120 : (ex > 120 ? 120 : ex);
            }
            const float s = std::ldexp(1.0f, ex), inv = std::ldexp(1.0f, -ex);
            Field E{storage, e.data()}, RT{storage, rt.data()};
            for_cells(g, active, n, [&](i64 i, i64, i64, i64) {
                RT.set(i, r[i] * inv);
                E.set(i, 0.0f);
            });
            for (int k = 0; k < per; ++k) {
                orc_rbgs(sx, sy, sz, storage, e.data(), rt.data(), b2s, active, n, 0, omega);
                orc_rbgs(sx, sy, sz, storage, e.data(), rt.data(), b2s, active, n, 1, omega);
            }
            for_cells(g, active, n, [&](i64 i, i64, i64, i64) { x[i] = fmaf(s, E.get(i), x[i]); });
        }
        sweeps += per;
        ++checks;
        orc_poisson_residual(sx, sy, sz, 0, x, b, b2s, active, n, r.data(), &r2, &b2);
    }
    out[0] = sweeps;
    out[1] = checks;
    out[2] = converged;
    out[3] = rel;
}

// FP64 conjugate gradient on the same operator (the FP64 oracle solution):
// x (double, slot-major, in/out) solves A x = b to ||r|| / ||b|| <= tol.
// Returns the iteration count.
int orc_cg64(i64 sx, i64 sy, i64 sz, double *x, const double *b, i64 n_slots, const i32 *b2s, const i64 *active,
             i64 n, double tol, int max_iter) {
    PGrid g(sx, sy, sz, b2s);
    const i64 N = n_slots * 4096;
    // compact unknowns and their 6 neighbours (-1: boundary value 0; the
    // mirrored neighbour at the domain edge is the cell itself)
    std::vector<i64> pos(N, -1), cell;
    for_cells(g, active, n, [&](i64 i, i64, i64, i64) {
        pos[i] = (i64)cell.size();
        cell.push_back(i);
    });
    const i64 M = (i64)cell.size();
    std::vector<i64> nb((size_t)M * 6);
    i64 k = 0;
    for_cells(g, active, n, [&](i64, i64 gx, i64 gy, i64 gz) {
        const i64 q[6] = {g.index(gx - 1, gy, gz), g.index(gx + 1, gy, gz), g.index(gx, gy - 1, gz),
                          g.index(gx, gy + 1, gz), g.index(gx, gy, gz - 1), g.index(gx, gy, gz + 1)};
        for (int d = 0; d < 6; ++d) nb[(size_t)k * 6 + d] = pos[q[d]];
        ++k;
    });
    std::vector<double> xv(M), r(M), p(M), ap(M);
    auto apply = [&](const std::vector<double> &v, std::vector<double> &out) {
#pragma omp parallel for schedule(static)
        for (i64 i = 0; i < M; ++i) {
            const i64 *q = &nb[(size_t)i * 6];
            double s = 0.0;
            for (int d = 0; d < 6; ++d) s += q[d] >= 0 ? v[q[d]] : 0.0;
            out[i] = 6.0 * v[i] - s;
        }
    };
    // dot products in fixed 4096-element chunks summed in order (deterministic)
    const i64 nch = (M + 4095) / 4096;
    std::vector<double> part(nch);
    auto dot = [&](const std::vector<double> &u, const std::vector<double> &v) {
#pragma omp parallel for schedule(static)
        for (i64 c = 0; c < nch; ++c) {
            double s = 0.0;
            for (i64 i = c * 4096; i < std::min(M, (c + 1) * 4096); ++i) s += u[i] * v[i];
            part[c] = s;
        }
        double s = 0.0;
        for (i64 c = 0; c < nch; ++c) s += part[c];
        return s;
    };
    std::vector<double> bv(M);
    for (i64 i = 0; i < M; ++i) {
        xv[i] = x[cell[i]];
        bv[i] = b[cell[i]];
    }
    apply(xv, ap);
    for (i64 i = 0; i < M; ++i) {
        r[i] = bv[i] - ap[i];
        p[i] = r[i];
    }
    double rr = dot(r, r);
    const double bb = dot(bv, bv);
    int it = 0;
    while (it < max_iter && std::sqrt(rr) > tol * std::sqrt(bb)) {
        apply(p, ap);
        const double alpha = rr / dot(p, ap);
#pragma omp parallel for schedule(static)
        for (i64 i = 0; i < M; ++i) {
            xv[i] += alpha * p[i];
            r[i] -= alpha * ap[i];
        }
        const double rr2 = dot(r, r);
        const double beta = rr2 / rr;
#pragma omp parallel for schedule(static)
        for (i64 i = 0; i < M; ++i) p[i] = r[i] + beta * p[i];
        rr = rr2;
        ++it;
    }
    for (i64 i = 0; i < M; ++i) x[cell[i]] = xv[i];
    return it;
}

}  // extern "C"

// CPU oracle, part 2: grid sampling (G2P), the reference MAC P2G splat, and
// the north-star transfer / solve methods that have no reference counterpart.
//
// TEST INFRASTRUCTURE ONLY (see oracle.cpp for the rules this file follows).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

typedef int64_t i64;
typedef int32_t i32;
typedef uint32_t u32;

namespace {

struct Grid {
    i64 sx, sy, sz, nx, ny, nz, nxy;
    const float *slab;
    const i32 *b2s;
    Grid(i64 sx_, i64 sy_, i64 sz_, const float *s, const i32 *b)
        : sx(sx_), sy(sy_), sz(sz_), slab(s), b2s(b) {
        nx = (sx + 15) >> 4;
        ny = (sy + 15) >> 4;
        nz = (sz + 15) >> 4;
        nxy = nx * ny;
    }
    const float *block_ptr(i64 bid) const { return slab + (i64)b2s[bid] * 4096; }
    float empty_value() const { return slab[0]; }
};

inline i32 reflect(i32 g, i32 max) {
    i32 period = 2 * (max + 1);
    i32 r = g % period;
    if (r < 0) r += period;
    return r > max ? period - 1 - r : r;
}

inline bool resolve_axis(i32 g, i32 max, int bc, i32 *out) {
    if (g >= 0 && g <= max) { *out = g; return true; }
    if (bc == 0) { *out = g < 0 ? 0 : max; return true; }
    if (bc == 1) { *out = reflect(g, max); return true; }
    return false;
}

// math/gather.rs:54-158 -- gather the (dmax+1)^3 window anchored at (ix0,iy0,iz0).
void gather_map(const Grid &g, i32 ix0, i32 iy0, i32 iz0, int bc, int dmax, float *out) {
    const int span = dmax + 1;
    if (ix0 >= 0 && iy0 >= 0 && iz0 >= 0 && (i64)ix0 + dmax < g.sx && (i64)iy0 + dmax < g.sy &&
        (i64)iz0 + dmax < g.sz) {
        int vx0 = ix0 & 15, vy0 = iy0 & 15, vz0 = iz0 & 15;
        if (vx0 + dmax < 16 && vy0 + dmax < 16 && vz0 + dmax < 16) {  // gather.rs:83-101
            i64 bid = (ix0 >> 4) + (i64)(iy0 >> 4) * g.nx + (i64)(iz0 >> 4) * g.nxy;
            const float *p = g.block_ptr(bid);
            int s = 0;
            for (int k = 0; k < span; ++k)
                for (int j = 0; j < span; ++j)
                    for (int i = 0; i < span; ++i) out[s++] = p[(vx0 + i) | ((vy0 + j) << 4) | ((vz0 + k) << 8)];
            return;
        }
    }
    // gather.rs:104-121: 2x2x2 block slots from the clamped anchor block.
    i64 bx0 = std::min(std::max((i64)(ix0 >> 4), (i64)0), g.nx - 1);
    i64 by0 = std::min(std::max((i64)(iy0 >> 4), (i64)0), g.ny - 1);
    i64 bz0 = std::min(std::max((i64)(iz0 >> 4), (i64)0), g.nz - 1);
    const float *slots[8];
    for (int sz = 0; sz < 2; ++sz)
        for (int sy = 0; sy < 2; ++sy)
            for (int sx = 0; sx < 2; ++sx) {
                i64 bxx = std::min(bx0 + sx, g.nx - 1), byy = std::min(by0 + sy, g.ny - 1),
                    bzz = std::min(bz0 + sz, g.nz - 1);
                slots[sx | (sy << 1) | (sz << 2)] = g.block_ptr(bxx + byy * g.nx + bzz * g.nxy);
            }
    // Where a Neumann reflection across a partial last block resolves outside
    // the table (slot index outside 0..7), the reference's bounds check panics
    // (gather.rs:143-146); the voxel's own block is read instead, which is the
    // table entry whenever the table covers it.
    auto slot_ptr = [&](i32 rx, i32 ry, i32 rz) -> const float * {
        const i64 ox = ((i64)rx >> 4) - bx0, oy = ((i64)ry >> 4) - by0, oz = ((i64)rz >> 4) - bz0;
        if (ox >= 0 && ox <= 1 && oy >= 0 && oy <= 1 && oz >= 0 && oz <= 1) return slots[ox | (oy << 1) | (oz << 2)];
        return g.block_ptr(((i64)rx >> 4) + ((i64)ry >> 4) * g.nx + ((i64)rz >> 4) * g.nxy);
    };
    const float empty = g.empty_value();
    int s = 0;
    for (int k = 0; k < span; ++k) {
        i32 rz = 0;
        bool okz = resolve_axis(iz0 + k, (i32)g.sz - 1, bc, &rz);
        for (int j = 0; j < span; ++j) {
            i32 ry = 0;
            bool oky = resolve_axis(iy0 + j, (i32)g.sy - 1, bc, &ry);
            for (int i = 0; i < span; ++i) {
                i32 rx = 0;
                bool okx = resolve_axis(ix0 + i, (i32)g.sx - 1, bc, &rx);
                if (okx && oky && okz) {
                    out[s] = slot_ptr(rx, ry, rz)[(rx & 15) | ((ry & 15) << 4) | ((rz & 15) << 8)];
                } else {
                    out[s] = empty;
                }
                ++s;
            }
        }
    }
}

inline float lerp(float a, float b, float t) { return fmaf(b - a, t, a); }  // sample.rs:194-196

float trilinear_value(const float *c, float u, float v, float w) {  // sample.rs:203-211
    float c00 = lerp(c[0], c[1], u), c10 = lerp(c[2], c[3], u);
    float c01 = lerp(c[4], c[5], u), c11 = lerp(c[6], c[7], u);
    float c0 = lerp(c00, c10, v), c1 = lerp(c01, c11, v);
    return lerp(c0, c1, w);
}

void trilinear_grad(const float *c, float u, float v, float w, float *g) {  // sample.rs:215-238
    float f0 = c[0], f1 = c[1], f2 = c[2], f3 = c[3], f4 = c[4], f5 = c[5], f6 = c[6], f7 = c[7];
    float f01234567 = f7 - f6 - f5 + f4 - f3 + f2 + f1 - f0;
    float f01234567u = f01234567 * u;
    float f01234567uf6420 = f01234567u + f6 - f4 - f2 + f0;
    float f5410 = f5 - f4 - f1 + f0;
    float f3210 = f3 - f2 - f1 + f0;
    g[0] = fmaf(f01234567 * v + f5410, w, f3210 * v + f1 - f0);
    g[1] = fmaf(f01234567uf6420, w, f3210 * u + f2 - f0);
    g[2] = fmaf(f01234567uf6420, v, f5410 * u + f4 - f0);
}

// math/bspline.rs:6-32 (Ruijters et al.)
void cubic_weights(float t, float *w) {
    float m = 1.0f - t;
    w[0] = (1.0f / 6.0f) * m * m * m;
    w[1] = 2.0f / 3.0f - 0.5f * t * t * (2.0f - t);
    w[2] = 2.0f / 3.0f - 0.5f * m * m * (1.0f + t);
    w[3] = (1.0f / 6.0f) * t * t * t;
}
void cubic_deriv_weights(float t, float *w) {
    float m = 1.0f - t;
    w[0] = -m * m / 2.0f;
    w[1] = 0.5f * t * t - (2.0f - t) * t;
    w[2] = m * (t + 1.0f) - 0.5f * m * m;
    w[3] = t * t / 2.0f;
}
void cubic_deriv2_weights(float t, float *w) {
    w[0] = 1.0f - t;
    w[1] = 3.0f * t - 2.0f;
    w[2] = 1.0f - 3.0f * t;
    w[3] = t;
}

// f32x4 reduce_sum is sequential: ((-0 + a0) + a1) + a2) + a3.
inline float dot4(const float *f, const float *w) {
    float s = -0.0f;
    s = s + f[0] * w[0];
    s = s + f[1] * w[1];
    s = s + f[2] * w[2];
    s = s + f[3] * w[3];
    return s;
}

// sample.rs:240-352 -- cubic value / gradient / Hessian over a 4^3 window.
void cubic_all(const float *win, float tx, float ty, float tz, float *val, float *grad, float *hess) {
    float wx[4], wpx[4], wppx[4], wy[4], wpy[4], wppy[4], wz[4], wpz[4], wppz[4];
    cubic_weights(tx, wx);
    cubic_deriv_weights(tx, wpx);
    cubic_deriv2_weights(tx, wppx);
    cubic_weights(ty, wy);
    cubic_deriv_weights(ty, wpy);
    cubic_deriv2_weights(ty, wppy);
    cubic_weights(tz, wz);
    cubic_deriv_weights(tz, wpz);
    cubic_deriv2_weights(tz, wppz);
    float v = 0, gx = 0, gy = 0, gz = 0, fxx = 0, fyy = 0, fzz = 0, fxy = 0, fxz = 0, fyz = 0;
    for (int k = 0; k < 4; ++k)
        for (int j = 0; j < 4; ++j) {
            const float *f = win + 16 * k + 4 * j;
            float sx = dot4(f, wx), sxp = dot4(f, wpx), sxpp = dot4(f, wppx);
            float wyz = wy[j] * wz[k];
            v = fmaf(sx, wyz, v);
            gx = fmaf(sxp, wyz, gx);
            gy = fmaf(sx, wpy[j] * wz[k], gy);
            gz = fmaf(sx, wy[j] * wpz[k], gz);
            fxx = fmaf(sxpp, wyz, fxx);
            fyy = fmaf(sx, wppy[j] * wz[k], fyy);
            fzz = fmaf(sx, wy[j] * wppz[k], fzz);
            fxy = fmaf(sxp, wpy[j] * wz[k], fxy);
            fxz = fmaf(sxp, wy[j] * wpz[k], fxz);
            fyz = fmaf(sx, wpy[j] * wpz[k], fyz);
        }
    *val = v;
    if (grad) { grad[0] = gx; grad[1] = gy; grad[2] = gz; }
    if (hess) { hess[0] = fxx; hess[1] = fyy; hess[2] = fzz; hess[3] = fxy; hess[4] = fxz; hess[5] = fyz; }
}

// Sample::sample (sample.rs:494-526) for one position.  interp 0 = Linear,
// 1 = CubicBSpline; align 0 = Corner, 1 = CellCentered (cubic ignores it).
float sample_one(const Grid &g, const float *p, int interp, int align, int bc, float *grad, float *hess) {
    if (interp == 0) {
        float x = p[0], y = p[1], z = p[2];
        if (align == 1) { x -= 0.5f; y -= 0.5f; z -= 0.5f; }
        float ix = floorf(x), iy = floorf(y), iz = floorf(z);
        float c[8];
        gather_map(g, (i32)ix, (i32)iy, (i32)iz, bc, 1, c);
        if (grad) trilinear_grad(c, x - ix, y - iy, z - iz, grad);
        return trilinear_value(c, x - ix, y - iy, z - iz);
    }
    float x = p[0] - 0.5f, y = p[1] - 0.5f, z = p[2] - 0.5f;  // sample.rs:465-467
    float xf = floorf(x), yf = floorf(y), zf = floorf(z);       // cubic_base sample.rs:471-474
    float win[64];
    gather_map(g, (i32)xf - 1, (i32)yf - 1, (i32)zf - 1, bc, 3, win);
    float v;
    cubic_all(win, x - xf, y - yf, z - zf, &v, grad, hess);
    return v;
}

inline float wcubic_falloff(float d2, float inv_r2) {
    float w = std::max(1.0f - d2 * inv_r2, 0.0f);
    return w * w * w;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// A15: Sampler::sample / gradient / hessian (math/sample.rs:489-580) over a
// batch of positions.  grad (n*3) and hess (n*6) may be null.
// ---------------------------------------------------------------------------
void orc_sample(i64 sx, i64 sy, i64 sz, const float *slab, const i32 *b2s, const float *pos, i64 n, int interp,
                int align, int bc, float *val, float *grad, float *hess) {
    Grid g(sx, sy, sz, slab, b2s);
    for (i64 i = 0; i < n; ++i) {
        float gr[3], he[6];
        float v = sample_one(g, pos + 3 * i, interp, align, bc, grad ? gr : nullptr, hess ? he : nullptr);
        if (val) val[i] = v;
        if (grad) std::memcpy(grad + 3 * i, gr, sizeof gr);
        if (hess) std::memcpy(hess + 6 * i, he, sizeof he);
    }
}

// ---------------------------------------------------------------------------
// MacSampler::sample (fluid/g2p.rs:26-32): face a sampled cell-centred at
// p - 1/2 e_a.  vel_slabs holds the three face-velocity slabs back to back,
// each n_slots*4096, over the shared b2s table.
// ---------------------------------------------------------------------------
void orc_mac_sample(i64 sx, i64 sy, i64 sz, const float *vel_slabs, i64 n_slots, const i32 *b2s, const float *pos,
                    i64 n, int interp, int bc, float *out) {
    for (int a = 0; a < 3; ++a) {
        Grid g(sx, sy, sz, vel_slabs + (i64)a * n_slots * 4096, b2s);
        for (i64 i = 0; i < n; ++i) {
            float q[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
            q[a] = q[a] - 0.5f;
            out[3 * i + a] = sample_one(g, q, interp, 1, bc, nullptr, nullptr);
        }
    }
}

// g2p_blend -- fluid/g2p.rs:42-45: (u_old + (u - u_tilde)) * a + u * (1 - a).
void orc_g2p_blend(const float *u_old, const float *u, const float *u_tilde, i64 n, float alpha, float *out) {
    for (i64 i = 0; i < 3 * n; ++i) {
        float delta = u[i] - u_tilde[i];
        out[i] = (u_old[i] + delta) * alpha + u[i] * (1.0f - alpha);
    }
}

// ---------------------------------------------------------------------------
// A12: particles_to_grid -> p2g::splat -> stage_and_commit ->
// normalize_velocity (fluid/p2g.rs:141-412, fluid/mac.rs:95-132) for r_p with
// margin m = ceil(r_p).  The caller passes the bucketed particles (positions,
// velocities, mass in block-major order), the occupied blocks and CSR starts,
// and a MAC grid whose 6 channels share `b2s` (active blocks zeroed, as
// MacGrid::zero_blocks).  slabs = [vel_x, vel_y, vel_z, mass_x, mass_y,
// mass_z], each n_slots*4096.  Blocks are processed per face direction in
// colour order (p2g.rs:186-216), each block staging its particles in order and
// committing its 27 regions (p2g.rs:335-411).  normalize: u = P/M, 0 if M<=0.
// With ref_commit_width = 1 and margin m < 4, the 4-lane commit of a dx = -1
// region (2 lanes wide) also adds staging columns m..3 -- interior cells of the
// centre block -- to the next row of the -x neighbour block; the reference
// tests only place particles in blocks with bx = 0, where that region is
// skipped (p2g.rs:367-369).
// ---------------------------------------------------------------------------
}  // extern "C"

namespace {
template <typename T>
void p2g_mac_impl(i64 sx, i64 sy, i64 sz, const float *pos, const float *vel, const float *mass,
                  const i64 *particle_blocks, const i64 *starts, i64 nblk, float r_p, T *slabs, i64 n_slots,
                  const i32 *b2s, const i64 *active, i64 n_active, int ref_commit_width, bool normalize) {
    const i64 nx = (sx + 15) >> 4, ny = (sy + 15) >> 4, nz = (sz + 15) >> 4, nxy = nx * ny;
    const int m = (int)ceilf(r_p);
    const int MSX = ((16 + 2 * m) + 7) & ~7;  // staging extent, multiple of 8 (p2g.rs:174-180)
    const float inv_r2 = 1.0f / (r_p * r_p);
    std::vector<std::pair<i64, i64>> by_color[8];
    for (i64 i = 0; i < nblk; ++i) {
        i64 bid = particle_blocks[i];
        i64 bx = bid % nx, by = (bid / nx) % ny, bz = bid / nxy;
        by_color[(bx & 1) | ((by & 1) << 1) | ((bz & 1) << 2)].push_back({i, bid});
    }
    std::vector<T> sm((size_t)MSX * MSX * MSX), sv((size_t)MSX * MSX * MSX);
    for (int a = 0; a < 3; ++a) {
        T *vel_g = slabs + (i64)a * n_slots * 4096;
        T *mass_g = slabs + (i64)(3 + a) * n_slots * 4096;
        const float ofx = a == 0 ? 1.0f : 0.5f, ofy = a == 1 ? 1.0f : 0.5f, ofz = a == 2 ? 1.0f : 0.5f;
        for (int c = 0; c < 8; ++c)
            for (auto &ib : by_color[c]) {
                const i64 i = ib.first, bid = ib.second;
                const i32 bx = (i32)(bid % nx), by = (i32)((bid / nx) % ny), bz = (i32)(bid / nxy);
                const i32 b0x = bx * 16, b0y = by * 16, b0z = bz * 16;
                std::fill(sm.begin(), sm.end(), T(0));
                std::fill(sv.begin(), sv.end(), T(0));
                const i32 x0s = b0x - m, y0s = b0y - m, z0s = b0z - m;
                const i32 window_hi = (16 + m) - 1;
                for (i64 pi = starts[i]; pi < starts[i + 1]; ++pi) {  // p2g.rs:261-333
                    const float *p = pos + 3 * pi;
                    const float scale_mass = mass[pi];
                    const float scale_mom = mass[pi] * vel[3 * pi + a];
                    const float px0 = p[0] - 0.5f, py0 = p[1] - 0.5f, pz0 = p[2] - 0.5f;
                    auto lo = [&](float v, i32 b0) { return std::min(std::max((i32)ceilf(v), b0 - m), b0 + window_hi); };
                    auto hi = [&](float v, i32 b0) { return std::min(std::max((i32)floorf(v), b0 - m), b0 + window_hi); };
                    i32 ix1 = lo(p[0] - r_p - ofx, b0x), ix2 = hi(p[0] + r_p - ofx, b0x);
                    i32 iy1 = lo(p[1] - r_p - ofy, b0y), iy2 = hi(p[1] + r_p - ofy, b0y);
                    i32 iz1 = lo(p[2] - r_p - ofz, b0z), iz2 = hi(p[2] + r_p - ofz, b0z);
                    for (i32 iz = iz1; iz <= iz2; ++iz) {
                        float dz = (float)iz - pz0;
                        float dz2 = dz * dz;
                        for (i32 iy = iy1; iy <= iy2; ++iy) {
                            float dy = (float)iy - py0;
                            float dyz = dz2 + dy * dy;
                            i64 row = ((i64)(iz - z0s) * MSX + (iy - y0s)) * MSX;
                            for (i32 ix = ix1; ix <= ix2; ++ix) {
                                i32 o = ix - x0s;
                                i32 c8 = o & ~7;
                                float dx = ((float)(c8 + x0s) + (float)(o - c8)) - px0;
                                float dist_sq = dx * dx + dyz;
                                float d_a = a == 0 ? dx : (a == 1 ? dy : dz);
                                float df2 = dist_sq + d_a + 0.25f;
                                float w = wcubic_falloff(df2, inv_r2);
                                sm[row + o] = sm[row + o] + (T)(w * scale_mass);  // the float product, summed in T
                                sv[row + o] = sv[row + o] + (T)(w * scale_mom);
                            }
                        }
                    }
                }
                // Commit the 27 regions (p2g.rs:342-411).
                for (int dz = -1; dz <= 1; ++dz)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dx = -1; dx <= 1; ++dx) {
                            auto rr = [&](int d, i32 *l, i32 *h) {
                                if (d == -1) { *l = 0; *h = m; }
                                else if (d == 0) { *l = m; *h = m + 16; }
                                else { *l = m + 16; *h = 2 * m + 16; }
                            };
                            i32 xlo, xhi, ylo, yhi, zlo, zhi;
                            rr(dx, &xlo, &xhi);
                            rr(dy, &ylo, &yhi);
                            rr(dz, &zlo, &zhi);
                            i32 gx0 = b0x + xlo - m, gy0 = b0y + ylo - m, gz0 = b0z + zlo - m;
                            i64 nbx = std::min(std::max((i64)bx + dx, (i64)0), nx - 1);
                            i64 nby = std::min(std::max((i64)by + dy, (i64)0), ny - 1);
                            i64 nbz = std::min(std::max((i64)bz + dz, (i64)0), nz - 1);
                            i64 nbid = nbx + nby * nx + nbz * nxy;
                            if (gx0 < 0 || gy0 < 0 || gz0 < 0) continue;
                            if (gx0 + (xhi - xlo) > sx || gy0 + (yhi - ylo) > sy || gz0 + (zhi - zlo) > sz) continue;
                            i32 s = b2s[nbid];
                            if (s < 2) continue;
                            T *md = mass_g + (i64)s * 4096, *vd = vel_g + (i64)s * 4096;
                            i32 lx0 = gx0 & 15, ly0 = gy0 & 15, lz0 = gz0 & 15;
                            // The reference's commit_sum adds W = 16 (interior) or 4
                            // (margin) lanes (p2g.rs:396-406) whatever the region
                            // width m; ref_commit_width = 1 reproduces that, 0 adds
                            // exactly the region's xhi - xlo lanes (the semantics of
                            // the naive oracle, p2g.rs:445-502).
                            i32 w = ref_commit_width ? (dx == 0 ? 16 : 4) : (xhi - xlo);
                            for (i32 szz = zlo; szz < zhi; ++szz)
                                for (i32 syy = ylo; syy < yhi; ++syy) {
                                    i64 srow = ((i64)szz * MSX + syy) * MSX + xlo;
                                    i64 drow = lx0 | ((ly0 + (syy - ylo)) << 4) | ((lz0 + (szz - zlo)) << 8);
                                    bool zm = true, zv = true;
                                    for (i32 k = 0; k < w; ++k) {
                                        zm = zm && sm[srow + k] == T(0);
                                        zv = zv && sv[srow + k] == T(0);
                                    }
                                    if (!zm)
                                        for (i32 k = 0; k < w; ++k) md[drow + k] = md[drow + k] + sm[srow + k];
                                    if (!zv)
                                        for (i32 k = 0; k < w; ++k) vd[drow + k] = vd[drow + k] + sv[srow + k];
                                }
                        }
            }
    }
    if (!normalize) return;
    // MacGrid::normalize_velocity (mac.rs:106-132).
    for (int a = 0; a < 3; ++a) {
        T *vel_g = slabs + (i64)a * n_slots * 4096;
        const T *mass_g = slabs + (i64)(3 + a) * n_slots * 4096;
        for (i64 k = 0; k < n_active; ++k) {
            i32 s = b2s[active[k]];
            if (s < 2) continue;
            for (int j = 0; j < 4096; ++j) {
                T mm = mass_g[(i64)s * 4096 + j];
                T &vv = vel_g[(i64)s * 4096 + j];
                vv = mm > T(0) ? vv / mm : T(0);
            }
        }
    }
}

}  // namespace

extern "C" {

void orc_p2g_mac(i64 sx, i64 sy, i64 sz, const float *pos, const float *vel, const float *mass,
                 const i64 *particle_blocks, const i64 *starts, i64 nblk, float r_p, float *slabs, i64 n_slots,
                 const i32 *b2s, const i64 *active, i64 n_active, int ref_commit_width) {
    p2g_mac_impl<float>(sx, sy, sz, pos, vel, mass, particle_blocks, starts, nblk, r_p, slabs, n_slots, b2s, active,
                        n_active, ref_commit_width, true);
}

// The same splat with every window and grid cell accumulated in float64: each
// contribution is the float32 product w * (m u_a) (or w * m) of the splat
// above, and the sums carry no float32 rounding (relative error <= k 2^-53
// for k terms).  Writes the unnormalised sums, slabs_f64 = [P_x, P_y, P_z,
// M_x, M_y, M_z] (each n_slots * 4096), exact-width commit.  Checks a splat
// whose float32 sums round differently from the reference's order (the
// reference's own order is unspecified, particles/sort.rs:234-237).
void orc_p2g_mac_f64(i64 sx, i64 sy, i64 sz, const float *pos, const float *vel, const float *mass,
                     const i64 *particle_blocks, const i64 *starts, i64 nblk, float r_p, double *slabs_f64,
                     i64 n_slots, const i32 *b2s, const i64 *active, i64 n_active) {
    p2g_mac_impl<double>(sx, sy, sz, pos, vel, mass, particle_blocks, starts, nblk, r_p, slabs_f64, n_slots, b2s,
                         active, n_active, 0, false);
}

// ---------------------------------------------------------------------------
// Naive Eq. 6 scalar P2G -- the reference's own equivalence oracle for the
// staged splat (fluid/p2g.rs:445-502): per particle, per direction, per face
// inside the clipped footprint, skip d2 > r^2, accumulate w*m and w*m*u.
// Dense output: 6 channels of sx*sy*sz (x fastest), normalized.
// ---------------------------------------------------------------------------
void orc_p2g_mac_naive(i64 sx, i64 sy, i64 sz, const float *pos, const float *vel, const float *mass, i64 n,
                       float r_p, float *dense) {
    const i64 nv = sx * sy * sz;
    std::memset(dense, 0, sizeof(float) * 6 * nv);
    const float inv_r2 = 1.0f / (r_p * r_p);
    for (i64 pi = 0; pi < n; ++pi) {
        const float *p = pos + 3 * pi;
        for (int a = 0; a < 3; ++a) {
            const float ofx = a == 0 ? 1.0f : 0.5f, ofy = a == 1 ? 1.0f : 0.5f, ofz = a == 2 ? 1.0f : 0.5f;
            i32 ix1 = std::max((i32)ceilf(p[0] - r_p - ofx), 0), ix2 = std::min((i32)floorf(p[0] + r_p - ofx), (i32)sx - 1);
            i32 iy1 = std::max((i32)ceilf(p[1] - r_p - ofy), 0), iy2 = std::min((i32)floorf(p[1] + r_p - ofy), (i32)sy - 1);
            i32 iz1 = std::max((i32)ceilf(p[2] - r_p - ofz), 0), iz2 = std::min((i32)floorf(p[2] + r_p - ofz), (i32)sz - 1);
            for (i32 k = iz1; k <= iz2; ++k)
                for (i32 j = iy1; j <= iy2; ++j)
                    for (i32 i = ix1; i <= ix2; ++i) {
                        float fx = (float)i + 0.5f + (a == 0 ? 0.5f : 0.0f);
                        float fy = (float)j + 0.5f + (a == 1 ? 0.5f : 0.0f);
                        float fz = (float)k + 0.5f + (a == 2 ? 0.5f : 0.0f);
                        float dx = fx - p[0], dy = fy - p[1], dz = fz - p[2];
                        float d2 = dx * dx + dy * dy + dz * dz;
                        if (d2 > r_p * r_p) continue;
                        float w = std::max(1.0f - d2 * inv_r2, 0.0f);
                        w = w * w * w;
                        i64 c = i + j * sx + k * sx * sy;
                        dense[(3 + a) * nv + c] = dense[(3 + a) * nv + c] + w * mass[pi];
                        dense[a * nv + c] = dense[a * nv + c] + w * mass[pi] * vel[3 * pi + a];
                    }
        }
    }
    for (int a = 0; a < 3; ++a)
        for (i64 c = 0; c < nv; ++c) {
            float mm = dense[(3 + a) * nv + c];
            dense[a * nv + c] = mm > 0.0f ? dense[a * nv + c] / mm : 0.0f;
        }
}

}  // extern "C"

// ===========================================================================
// Density min-splat -- particles/splat.rs:92-319 with D = Density (u16,
// channel.rs:303-375).  Particles are grouped by the block of their truncated
// position; each block's group is staged into a (16 + 2m)^3 window (m =
// ceil(r_scan)) filled with 65535: every voxel ix in [ceil(p - r - 0.5),
// floor(p + r - 0.5)] per axis (clipped to the window) with
//   dx = ix - (p.x - 0.5), dz2 = dz*dz, dyz = dz2 + dy*dy, d2 = dx*dx + dyz,
// d2 <= r^2 takes min(staging, (u16)round_half_away(d2 * (1/r^2) * 65535))
// (splat.rs:179-225, stage_chunk channel.rs:353-374).  The commit splits the
// window into 3x3x3 regions (interior 16, margins m); a region is skipped
// when any of its voxels lies outside the domain or its target block is not a
// value block (splat.rs:240-292); otherwise each region row commits
// dst = min(dst, staging) (commit_chunk, channel.rs:337-351).
// ref_commit_width = 1 reproduces the reference's row width for margin rows
// (4 lanes, splat.rs:306-309), which spills past the region when m < 4 (into
// the next row of the -x neighbour, as MAC P2G's margin commit); 0 commits
// each region over its own width.  For 3 < r_scan <= 4 (the tests' and the
// demo's r_scan = 4) the two agree.
// ===========================================================================
namespace {

uint16_t quant_dsq(float d2, float inv) {
    const float v = roundf((d2 * inv) * 65535.0f);  // f32::round: half away from zero
    return (uint16_t)(int32_t)v;
}

}  // namespace

extern "C" void orc_density_splat(i64 sx, i64 sy, i64 sz, const i32 *b2s, uint16_t *slab, const float *pos, i64 n,
                                  float r_scan, int ref_commit_width) {
    const i64 nx = (sx + 15) >> 4, ny = (sy + 15) >> 4, nxy = nx * ny;
    const int m = (int)ceilf(r_scan);
    const int W = 16 + 2 * m;
    const int MSX = std::max(24, (W + 7) & ~7);  // staging row pitch (the reference's MSX, a multiple of 8)
    const float dsq_max = r_scan * r_scan, inv = 1.0f / dsq_max;
    // group particles by block, in bid order (stable)
    std::vector<i64> bid((size_t)n);
    for (i64 i = 0; i < n; ++i) {
        auto tr = [](float v) { return (i64)(v > 0.0f ? (uint64_t)v : 0); };  // `as usize` (saturating)
        const i64 x = tr(pos[3 * i]), y = tr(pos[3 * i + 1]), z = tr(pos[3 * i + 2]);
        bid[i] = (x >> 4) + (y >> 4) * nx + (z >> 4) * nxy;
    }
    std::vector<i64> order((size_t)n);
    for (i64 i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](i64 a, i64 b) { return bid[a] < bid[b]; });
    std::vector<uint16_t> st((size_t)MSX * MSX * MSX + 8);
    for (i64 k0 = 0; k0 < n;) {
        i64 k1 = k0;
        while (k1 < n && bid[order[k1]] == bid[order[k0]]) ++k1;
        const i64 b = bid[order[k0]];
        const int bx = (int)(b % nx), by = (int)((b / nx) % ny), bz = (int)(b / nxy);
        const int b0x = bx * 16, b0y = by * 16, b0z = bz * 16, hi = 16 + m - 1;
        std::fill(st.begin(), st.end(), (uint16_t)65535);
        for (i64 k = k0; k < k1; ++k) {
            const float *p = pos + 3 * order[k];
            const float px0 = p[0] - 0.5f, py0 = p[1] - 0.5f, pz0 = p[2] - 0.5f;
            auto lo_of = [&](float c, int b0) { return std::min(std::max((int)ceilf((c - r_scan) - 0.5f), b0 - m), b0 + hi); };
            auto hi_of = [&](float c, int b0) { return std::min(std::max((int)floorf((c + r_scan) - 0.5f), b0 - m), b0 + hi); };
            const int ix1 = lo_of(p[0], b0x), ix2 = hi_of(p[0], b0x), iy1 = lo_of(p[1], b0y), iy2 = hi_of(p[1], b0y),
                      iz1 = lo_of(p[2], b0z), iz2 = hi_of(p[2], b0z);
            for (int iz = iz1; iz <= iz2; ++iz) {
                const float dz = (float)iz - pz0, dz2 = dz * dz;
                for (int iy = iy1; iy <= iy2; ++iy) {
                    const float dy = (float)iy - py0, dyz = dz2 + dy * dy;
                    uint16_t *row = &st[((size_t)(iz - b0z + m) * MSX + (iy - b0y + m)) * MSX];
                    for (int ix = ix1; ix <= ix2; ++ix) {
                        const float dx = (float)ix - px0, d2 = dx * dx + dyz;
                        if (d2 <= dsq_max) {
                            uint16_t &s = row[ix - b0x + m];
                            s = std::min(s, quant_dsq(d2, inv));
                        }
                    }
                }
            }
        }
        // commit
        for (int rz = -1; rz <= 1; ++rz)
            for (int ry = -1; ry <= 1; ++ry)
                for (int rx = -1; rx <= 1; ++rx) {
                    auto range = [&](int d, int *lo, int *hi2) {
                        *lo = d < 0 ? 0 : (d == 0 ? m : m + 16);
                        *hi2 = d < 0 ? m : (d == 0 ? m + 16 : 2 * m + 16);
                    };
                    int xlo, xhi, ylo, yhi, zlo, zhi;
                    range(rx, &xlo, &xhi);
                    range(ry, &ylo, &yhi);
                    range(rz, &zlo, &zhi);
                    const int gx0 = b0x + xlo - m, gy0 = b0y + ylo - m, gz0 = b0z + zlo - m;
                    if (gx0 < 0 || gy0 < 0 || gz0 < 0) continue;
                    if (gx0 + (xhi - xlo) > sx || gy0 + (yhi - ylo) > sy || gz0 + (zhi - zlo) > sz) continue;
                    const i64 nb = (gx0 >> 4) + (i64)(gy0 >> 4) * nx + (i64)(gz0 >> 4) * nxy;
                    const i32 slot = b2s[nb];
                    if (slot < 2) continue;
                    const int w = (ref_commit_width && rx != 0) ? 4 : (xhi - xlo);
                    for (int sz_ = zlo; sz_ < zhi; ++sz_)
                        for (int sy_ = ylo; sy_ < yhi; ++sy_) {
                            const uint16_t *src = &st[((size_t)sz_ * MSX + sy_) * MSX + xlo];
                            const i64 d = (i64)slot * 4096 + ((gx0 & 15) | (((gy0 + sy_ - ylo) & 15) << 4) |
                                                              (((gz0 + sz_ - zlo) & 15) << 8));
                            for (int k = 0; k < w; ++k) slab[d + k] = std::min(slab[d + k], src[k]);
                        }
                }
        k0 = k1;
    }
}

// ===========================================================================
// MacGrid::divergence -- fluid/mac.rs:137-171: for the cells of the `active`
// blocks of dst, ((u(x) - u(x-1)) + (v(y) - v(y-1))) + (w(z) - w(z-1)), a face
// at index 0 on its axis reading 0; velocity voxels read through each
// channel's own bid -> slot table (absent / full blocks give the dummies).
// The reference stores cell k of a block's clipped extent at vid = k (a
// running index, mac.rs:155-166), which is the cell's layout index only for
// whole 16^3 blocks; ref_packed_vid = 1 reproduces that, 0 stores every cell
// at its layout index x | y << 4 | z << 8.
// ===========================================================================
extern "C" void orc_mac_divergence(i64 sx, i64 sy, i64 sz, const float *u, const i32 *ub2s, const float *v,
                                   const i32 *vb2s, const float *w, const i32 *wb2s, float *dst, const i32 *db2s,
                                   const i64 *active, i64 n, int ref_packed_vid) {
    const i64 nx = (sx + 15) >> 4, ny = (sy + 15) >> 4, nxy = nx * ny;
    auto at = [&](const float *s, const i32 *b2s, i64 x, i64 y, i64 z) {
        return s[(i64)b2s[(x >> 4) + (y >> 4) * nx + (z >> 4) * nxy] * 4096 + ((x & 15) | ((y & 15) << 4) | ((z & 15) << 8))];
    };
    for (i64 k = 0; k < n; ++k) {
        const i64 bid = active[k];
        const i32 slot = db2s[bid];
        if (slot < 2) continue;
        const i64 gx0 = (bid % nx) * 16, gy0 = ((bid / nx) % ny) * 16, gz0 = (bid / nxy) * 16;
        const i64 gx1 = std::min(gx0 + 16, sx), gy1 = std::min(gy0 + 16, sy), gz1 = std::min(gz0 + 16, sz);
        i64 run = 0;
        for (i64 z = gz0; z < gz1; ++z)
            for (i64 y = gy0; y < gy1; ++y)
                for (i64 x = gx0; x < gx1; ++x) {
                    const float ux1 = at(u, ub2s, x, y, z), ux0 = x == 0 ? 0.0f : at(u, ub2s, x - 1, y, z);
                    const float vy1 = at(v, vb2s, x, y, z), vy0 = y == 0 ? 0.0f : at(v, vb2s, x, y - 1, z);
                    const float wz1 = at(w, wb2s, x, y, z), wz0 = z == 0 ? 0.0f : at(w, wb2s, x, y, z - 1);
                    const i64 vid = ref_packed_vid ? run : ((x & 15) | ((y & 15) << 4) | ((z & 15) << 8));
                    dst[(i64)slot * 4096 + vid] = ((ux1 - ux0) + (vy1 - vy0)) + (wz1 - wz0);
                    ++run;
                }
    }
}

// ===========================================================================
// Pressure projection of the face velocities (no reference counterpart: the
// reference stops at the divergence, mac.rs:134-171, the projection being its
// planned step 18).  Face a of cell c (the face between c and c + e_a, the
// layout `divergence` reads) takes u_a(c) - scale * (p(c + e_a) - p(c)); a
// neighbour outside the domain is the cell itself (zero gradient, the
// Poisson operator's mirror), a neighbour in an absent block reads p's slot 0.
// With A p = -div (or scale = -1 and A p = div) the new divergence is the
// solve's residual.  Every cell of the listed blocks' clipped extent, in
// place on each channel (the three channels' own bid -> slot tables).
// ===========================================================================
// The projection from a compensated pressure p + plo (plo over p's slot
// layout, or NULL): the gradient is (p(c + e_a) - p(c)) + (plo(c + e_a) -
// plo(c)).
extern "C" void orc_mac_subtract_gradient_pair(i64 sx, i64 sy, i64 sz, float *u, const i32 *ub2s, float *v,
                                               const i32 *vb2s, float *w, const i32 *wb2s, const float *p,
                                               const float *plo, const i32 *pb2s, const i64 *active, i64 n,
                                               float scale);

extern "C" void orc_mac_subtract_gradient(i64 sx, i64 sy, i64 sz, float *u, const i32 *ub2s, float *v,
                                          const i32 *vb2s, float *w, const i32 *wb2s, const float *p,
                                          const i32 *pb2s, const i64 *active, i64 n, float scale) {
    orc_mac_subtract_gradient_pair(sx, sy, sz, u, ub2s, v, vb2s, w, wb2s, p, nullptr, pb2s, active, n, scale);
}

extern "C" void orc_mac_subtract_gradient_pair(i64 sx, i64 sy, i64 sz, float *u, const i32 *ub2s, float *v,
                                               const i32 *vb2s, float *w, const i32 *wb2s, const float *p,
                                               const float *plo, const i32 *pb2s, const i64 *active, i64 n,
                                               float scale) {
    const i64 nx = (sx + 15) >> 4, ny = (sy + 15) >> 4, nxy = nx * ny;
    auto idx = [&](const i32 *b2s, i64 x, i64 y, i64 z) {
        return (i64)b2s[(x >> 4) + (y >> 4) * nx + (z >> 4) * nxy] * 4096 + ((x & 15) | ((y & 15) << 4) | ((z & 15) << 8));
    };
    float *ch[3] = {u, v, w};
    const i32 *cb[3] = {ub2s, vb2s, wb2s};
    for (i64 k = 0; k < n; ++k) {
        const i64 bid = active[k];
        const i64 gx0 = (bid % nx) * 16, gy0 = ((bid / nx) % ny) * 16, gz0 = (bid / nxy) * 16;
        const i64 gx1 = std::min(gx0 + 16, sx), gy1 = std::min(gy0 + 16, sy), gz1 = std::min(gz0 + 16, sz);
        for (i64 z = gz0; z < gz1; ++z)
            for (i64 y = gy0; y < gy1; ++y)
                for (i64 x = gx0; x < gx1; ++x) {
                    const float pc = p[idx(pb2s, x, y, z)];
                    const float pn[3] = {x + 1 < sx ? p[idx(pb2s, x + 1, y, z)] : pc,
                                         y + 1 < sy ? p[idx(pb2s, x, y + 1, z)] : pc,
                                         z + 1 < sz ? p[idx(pb2s, x, y, z + 1)] : pc};
                    float lc = 0.0f, ln[3] = {0.0f, 0.0f, 0.0f};
                    if (plo) {
                        lc = plo[idx(pb2s, x, y, z)];
                        ln[0] = x + 1 < sx ? plo[idx(pb2s, x + 1, y, z)] : lc;
                        ln[1] = y + 1 < sy ? plo[idx(pb2s, x, y + 1, z)] : lc;
                        ln[2] = z + 1 < sz ? plo[idx(pb2s, x, y, z + 1)] : lc;
                    }
                    for (int a = 0; a < 3; ++a) {
                        if (cb[a][bid] < 2) continue;
                        float &f = ch[a][idx(cb[a], x, y, z)];
                        float g = pn[a] - pc;
                        if (plo) g = g + (ln[a] - lc);
                        const float t = scale * g;
                        f = f - t;
                    }
                }
    }
}

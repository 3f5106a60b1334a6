// CPU oracle, part 5: the surface-reconstruction tail -- instance placement
// (particles/sort.rs `place`) and the distance-to-density finalize
// (particles/finalize.rs) -- and the per-block topology flags of a multires
// refinement map (multires/refinement.rs `compute_block_topology`).
//
// TEST INFRASTRUCTURE ONLY (see oracle.cpp for the rules this file follows):
// plain scalar C++ in the reference's operation order, built with
// -ffp-contract=off so no multiply-add is fused.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

typedef int64_t i64;
typedef int32_t i32;

namespace {

// align24 -- particles/mod.rs:180-183 (the C++ PARTICLE_QUANT_POS bound).
float align24(double x) {
    const double scale = 16777215.0 / 32768.0;
    return (float)(std::floor(scale * x + 0.5) / scale);
}

// Block flag bits -- multires/blockinfo.rs:39-54.
constexpr uint16_t kDomBorder = 1u << 1, kCoarseFine = 1u << 2, kFineCoarse = 1u << 3, kExists = 1u << 12;

}  // namespace

extern "C" {

// Domain bounds of the placement check -- DomainBounds::new,
// particles/mod.rs:185-194: out[0..3] = min, out[3..6] = max.
void orc_domain_bounds(i64 sx, i64 sy, i64 sz, float *out) {
    const i64 s[3] = {sx, sy, sz};
    for (int k = 0; k < 3; ++k) {
        out[k] = align24(0.005);
        out[3 + k] = align24((double)s[k] - 0.005);
    }
}

// place -- particles/sort.rs:31-143.  Instance i (0 <= i < n_instances, 0 =
// one per base point) has origin 0.2 smax + 0.6 smax base_scale (b_i - bbox)
// and places every base point j at origin + inst_scale (b_j - bbox); a
// position inside the domain bounds whose truncation lies in the grid is
// kept, with its centre block id.  Output order: instance-major, base point
// minor (the order rayon's fold/reduce concatenates).  Returns the number of
// kept particles, or -1 if it would exceed `cap` (outputs then partial).
i64 orc_surface_place(const float *base, i64 n_base, const float *bbox_min, float span_max, i64 sx, i64 sy, i64 sz,
                      i64 n_instances, float instance_scale_factor, float *pos_out, i64 *bid_out, i64 cap) {
    const i64 nx = (sx + 15) >> 4, ny = (sy + 15) >> 4, nxy = nx * ny;
    const float base_scale = 1.0f / span_max;
    const float smin = (float)std::min(sx, std::min(sy, sz));
    const float smax = (float)std::max(sx, std::max(sy, sz));
    const float scale = smin * instance_scale_factor;
    const float inst_scale = scale * base_scale;
    float bounds[6];
    orc_domain_bounds(sx, sy, sz, bounds);
    const i64 s[3] = {sx, sy, sz};
    const i64 n_inst = n_instances == 0 ? n_base : n_instances;
    i64 k = 0;
    for (i64 i = 0; i < n_inst; ++i) {
        float origin[3];
        for (int a = 0; a < 3; ++a) {
            const float t0 = 0.2f * smax;
            float t1 = 0.6f * smax;
            t1 = t1 * base_scale;
            t1 = t1 * (base[3 * i + a] - bbox_min[a]);
            origin[a] = t0 + t1;
        }
        for (i64 j = 0; j < n_base; ++j) {
            float p[3];
            bool in = true;
            for (int a = 0; a < 3; ++a) {
                p[a] = origin[a] + inst_scale * (base[3 * j + a] - bbox_min[a]);
                in = in && p[a] >= bounds[a] && p[a] <= bounds[3 + a];
            }
            if (!in) continue;
            i32 ip[3];
            bool ok = true;
            for (int a = 0; a < 3; ++a) {
                ip[a] = (i32)p[a];
                ok = ok && ip[a] >= 0 && ip[a] < (i32)s[a];
            }
            if (!ok) continue;
            if (k >= cap) return -1;
            for (int a = 0; a < 3; ++a) pos_out[3 * k + a] = p[a];
            bid_out[k] = (i64)(ip[0] >> 4) + (i64)(ip[1] >> 4) * nx + (i64)(ip[2] >> 4) * nxy;
            ++k;
        }
    }
    return k;
}

// finalize -- particles/finalize.rs:35-71 over a Density (u16) slab: ratio =
// u * (1/65535); touched = ratio < 0.999; d = sqrt(ratio r_scan^2) - r_p;
// f = 1 - clamp((d + r_p) / (2 r_p + nb), 0, 1); store round(f * 65535) half
// away from zero (Density::quant_chunk, channel.rs:329-335), 0 if untouched.
// Blocks without a value slot are skipped.
void orc_density_finalize(const i32 *b2s, uint16_t *slab, const i64 *active, i64 n, float r_particle,
                          float nb_dist) {
    const float r_scan = r_particle + nb_dist;
    const float dist_sq_max = r_scan * r_scan;
    const float denom = 2.0f * r_particle + nb_dist;
    const float inv = 1.0f / 65535.0f;
    for (i64 a = 0; a < n; ++a) {
        const i32 slot = b2s[active[a]];
        if (slot < 2) continue;
        uint16_t *d = slab + (i64)slot * 4096;
        for (int v = 0; v < 4096; ++v) {
            const float ratio = (float)d[v] * inv;
            const bool touched = ratio < 0.999f;
            const float dist_sq = ratio * dist_sq_max;
            const float dist = std::sqrt(dist_sq) - r_particle;
            float t = (dist + r_particle) / denom;
            t = std::min(std::max(t, 0.0f), 1.0f);
            const float f = 1.0f - t;
            const float out = touched ? f : 0.0f;
            d[v] = (uint16_t)(i32)std::round(out * 65535.0f);
        }
    }
}

// compute_block_topology -- multires/refinement.rs:124-208: flags = user
// flags | DOM_BORDER on the block-grid faces | FINE_COARSE if a 26-neighbour
// is coarser | COARSE_FINE if one is finer.  fc_bids receives the
// FINE_COARSE blocks of level 0, then level 1, ... (ascending within a
// level), fc_counts[l] their counts.  Returns the EXISTS count.
i64 orc_block_topology(i64 nx, i64 ny, i64 nz, const uint8_t *levels, int n_levels, const uint16_t *user_flags,
                       uint16_t *flags_out, i64 *fc_bids, i64 *fc_counts) {
    const i64 nxy = nx * ny, nb = nxy * nz;
    std::vector<std::vector<i64>> lists((size_t)n_levels);
    i64 n_act = 0;
    for (i64 bid = 0; bid < nb; ++bid) {
        const int level = levels[bid];
        uint16_t f = user_flags ? user_flags[bid] : 0;
        const i64 bx = bid % nx, by = (bid / nx) % ny, bz = bid / nxy;
        if (bx == 0 || bx == nx - 1 || by == 0 || by == ny - 1 || bz == 0 || bz == nz - 1) f |= kDomBorder;
        int lmax = 0, lmin = n_levels - 1;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    if (!dx && !dy && !dz) continue;
                    const i64 x = bx + dx, y = by + dy, z = bz + dz;
                    if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) continue;
                    const int l2 = levels[x + y * nx + z * nxy];
                    lmax = std::max(lmax, l2);
                    lmin = std::min(lmin, l2);
                }
        if (lmax > level) f |= kFineCoarse;
        if (lmin < level) f |= kCoarseFine;
        flags_out[bid] = f;
        if (f & kFineCoarse) lists[(size_t)level].push_back(bid);
        if (f & kExists) ++n_act;
    }
    i64 k = 0;
    for (int l = 0; l < n_levels; ++l) {
        fc_counts[l] = (i64)lists[(size_t)l].size();
        for (i64 b : lists[(size_t)l]) fc_bids[k++] = b;
    }
    return n_act;
}

}  // extern "C"

// Particle <-> grid transfers.
//
//  * Sampler (G2P of the reference): Sample::sample (math/sample.rs:489-526)
//    with gather_map (math/gather.rs:54-158), per particle, plus the MAC
//    face sampler (fluid/g2p.rs:26-32) and the FLIP/PIC blend (g2p.rs:42-45).
//  * MAC P2G of the reference (fluid/p2g.rs:141-412): particles bucketed by
//    block; per face direction and colour one launch, each workgroup stages
//    its block's (16+2m)^3 window in LDS (mass + momentum) and commits the 27
//    regions to the grid; blocks of one colour never overlap, so the commit
//    needs no global atomics.  Then u = P/M (mac.rs:106-132).
//  * B-spline P2G / G2P (trilinear, quadratic): particles Morton-sorted by
//    voxel; every cell gathers its 27 neighbour voxels in a fixed order, so
//    the result is deterministic and needs neither atomics nor colouring.
#include <algorithm>
#include <climits>
#include <vector>

#include "primitives.hpp"

namespace msbg {

unsigned grid_for(int64_t n, int threads);
int run_lengths(const uint32_t *keys, int64_t n, int32_t *blocks, int32_t *starts, int64_t *n_occ);

// ---------------------------------------------------------------------------
// Reference sampler.
// ---------------------------------------------------------------------------
__device__ __forceinline__ const float *block_ptr(const GridView &g, int64_t bid) {
    return g.slab + (int64_t)g.b2s[bid] * 4096;
}

// gather_map (math/gather.rs:54-158).
__device__ void gather_map_dev(const GridView &g, int ix0, int iy0, int iz0, int bc, int dmax, float *out) {
    const int span = dmax + 1;
    if (ix0 >= 0 && iy0 >= 0 && iz0 >= 0 && ix0 + dmax < g.sx && iy0 + dmax < g.sy && iz0 + dmax < g.sz) {
        const int vx0 = ix0 & 15, vy0 = iy0 & 15, vz0 = iz0 & 15;
        if (vx0 + dmax < 16 && vy0 + dmax < 16 && vz0 + dmax < 16) {
            const float *p = block_ptr(g, (ix0 >> 4) + (int64_t)(iy0 >> 4) * g.nx + (int64_t)(iz0 >> 4) * g.nx * g.ny);
            int s = 0;
            for (int k = 0; k < span; ++k)
                for (int j = 0; j < span; ++j)
                    for (int i = 0; i < span; ++i) out[s++] = p[(vx0 + i) | ((vy0 + j) << 4) | ((vz0 + k) << 8)];
            return;
        }
    }
    // gather.rs:104-121 reads the resolved voxel through a 2x2x2 table of
    // blocks anchored at the clamped window corner.  The voxel's own block is
    // looked up here instead: the same block whenever the table covers it, and
    // in bounds where it does not (a Neumann reflection across a partial last
    // block), where the reference's slot bounds check panics (gather.rs:143-146).
    const float empty = g.slab[0];
    if (dmax == 1) {
        // The 8 cells' slots and values are each loaded in one batch (a
        // window across a block face reads up to 8 blocks): every lookup is
        // issued before any is used.  Same cells, same values as below.
        int64_t bid[8];
        int vid[8];
        uint32_t ok = 0;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            int rx, ry, rz;
            const bool in = resolve_axis(ix0 + (c & 1), g.sx - 1, bc, rx) &&
                            resolve_axis(iy0 + ((c >> 1) & 1), g.sy - 1, bc, ry) &&
                            resolve_axis(iz0 + (c >> 2), g.sz - 1, bc, rz);
            bid[c] = in ? (rx >> 4) + (int64_t)(ry >> 4) * g.nx + (int64_t)(rz >> 4) * g.nx * g.ny : 0;
            vid[c] = in ? (rx & 15) | ((ry & 15) << 4) | ((rz & 15) << 8) : 0;
            ok |= (uint32_t)in << c;
        }
        int slot[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) slot[c] = g.b2s[bid[c]];
#pragma unroll
        for (int c = 0; c < 8; ++c) out[c] = g.slab[(int64_t)slot[c] * 4096 + vid[c]];
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (!((ok >> c) & 1u)) out[c] = empty;
        return;
    }
    int s = 0;
    for (int k = 0; k < span; ++k) {
        int rz;
        const bool okz = resolve_axis(iz0 + k, g.sz - 1, bc, rz);
        for (int j = 0; j < span; ++j) {
            int ry;
            const bool oky = resolve_axis(iy0 + j, g.sy - 1, bc, ry);
            for (int i = 0; i < span; ++i) {
                int rx;
                const bool okx = resolve_axis(ix0 + i, g.sx - 1, bc, rx);
                if (okx && oky && okz) {
                    const float *p = block_ptr(g, (rx >> 4) + (int64_t)(ry >> 4) * g.nx + (int64_t)(rz >> 4) * g.nx * g.ny);
                    out[s] = p[(rx & 15) | ((ry & 15) << 4) | ((rz & 15) << 8)];
                } else {
                    out[s] = empty;
                }
                ++s;
            }
        }
    }
}

__device__ __forceinline__ float lerp_f(float a, float b, float t) { return __builtin_fmaf(b - a, t, a); }

__device__ __forceinline__ float dot4_seq(const float *f, const float *w) {  // f32x4 reduce_sum, sequential
    float s = -0.0f;
    s = s + f[0] * w[0];
    s = s + f[1] * w[1];
    s = s + f[2] * w[2];
    s = s + f[3] * w[3];
    return s;
}

__device__ __forceinline__ void cubic_weights_dev(float t, float *w) {  // math/bspline.rs:6-14
    const float m = 1.0f - t;
    w[0] = (1.0f / 6.0f) * m * m * m;
    w[1] = 2.0f / 3.0f - 0.5f * t * t * (2.0f - t);
    w[2] = 2.0f / 3.0f - 0.5f * m * m * (1.0f + t);
    w[3] = (1.0f / 6.0f) * t * t * t;
}

// Sample::sample (math/sample.rs:494-526).  IP 0 Linear, 1 CubicBSpline;
// align 0 Corner, 1 CellCentered (cubic is always cell-centred).
template <int IP>
__device__ __forceinline__ float sample_ip(const GridView &g, float px, float py, float pz, int align, int bc) {
    if constexpr (IP == 0) {
        float x = px, y = py, z = pz;
        if (align == 1) {
            x = x - 0.5f;
            y = y - 0.5f;
            z = z - 0.5f;
        }
        const float ix = floorf(x), iy = floorf(y), iz = floorf(z);
        float c[8];
        gather_map_dev(g, (int)ix, (int)iy, (int)iz, bc, 1, c);
        const float u = x - ix, v = y - iy, w = z - iz;
        const float c00 = lerp_f(c[0], c[1], u), c10 = lerp_f(c[2], c[3], u);
        const float c01 = lerp_f(c[4], c[5], u), c11 = lerp_f(c[6], c[7], u);
        return lerp_f(lerp_f(c00, c10, v), lerp_f(c01, c11, v), w);
    } else {
    const float x = px - 0.5f, y = py - 0.5f, z = pz - 0.5f;
    const float xf = floorf(x), yf = floorf(y), zf = floorf(z);
    float win[64];
    gather_map_dev(g, (int)xf - 1, (int)yf - 1, (int)zf - 1, bc, 3, win);
    float wx[4], wy[4], wz[4];
    cubic_weights_dev(x - xf, wx);
    cubic_weights_dev(y - yf, wy);
    cubic_weights_dev(z - zf, wz);
    float val = 0.0f;
    for (int k = 0; k < 4; ++k)
        for (int j = 0; j < 4; ++j) val = __builtin_fmaf(dot4_seq(win + 16 * k + 4 * j, wx), wy[j] * wz[k], val);
    return val;
    }
}

__device__ float sample_dev(const GridView &g, float px, float py, float pz, int interp, int align, int bc) {
    return interp == 0 ? sample_ip<0>(g, px, py, pz, align, bc) : sample_ip<1>(g, px, py, pz, align, bc);
}

// out[i * stride] = sample at p_i - 1/2 e_axis (axis < 0: no offset).
template <int IP>
__global__ void sample_kernel(GridView g, const float *__restrict__ pos, int64_t n, int align, int bc, int axis,
                              float *__restrict__ out, int stride) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float q[3] = {axis == 0 ? pos[3 * i] - 0.5f : pos[3 * i], axis == 1 ? pos[3 * i + 1] - 0.5f : pos[3 * i + 1],
                            axis == 2 ? pos[3 * i + 2] - 0.5f : pos[3 * i + 2]};
        out[i * stride] = sample_ip<IP>(g, q[0], q[1], q[2], align, bc);
    }
}

void launch_sample(const GridView &g, const float *pos, int64_t n, int interp, int align, int bc, int axis,
                   float *out, int stride) {
    if (interp == 0)
        hipLaunchKernelGGL(sample_kernel<0>, dim3(grid_for(n, 256)), dim3(256), 0, stream(), g, pos, n, align, bc, axis,
                           out, stride);
    else
        hipLaunchKernelGGL(sample_kernel<1>, dim3(grid_for(n, 256)), dim3(256), 0, stream(), g, pos, n, align, bc, axis,
                           out, stride);
}

__global__ void blend_kernel(const float *__restrict__ u_old, const float *__restrict__ u,
                             const float *__restrict__ u_tilde, int64_t n3, float alpha, float *__restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (int64_t)gridDim.x * blockDim.x) {
        const float delta = u[i] - u_tilde[i];  // g2p.rs:43-44
        out[i] = (u_old[i] + delta) * alpha + u[i] * (1.0f - alpha);
    }
}

// G2P of a FLIP step, per particle (fluid/g2p.rs:11-45): the staggered
// samples of u and u* (MacSampler::sample: face a at p - 1/2 e_a) and the
// FLIP/PIC blend, with the arithmetic of sample_kernel and blend_kernel.
struct MacViews {
    GridView v[3];
};
template <int IP>
__global__ void g2p_flip_kernel(MacViews u, MacViews ut, const float *__restrict__ pos,
                                const float *__restrict__ u_old, int64_t n, int bc, float alpha,
                                float *__restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float p[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
#pragma unroll 1
        for (int a = 0; a < 3; ++a) {  // one axis at a time keeps the live registers of one sample pair
            const float q[3] = {a == 0 ? p[0] - 0.5f : p[0], a == 1 ? p[1] - 0.5f : p[1], a == 2 ? p[2] - 0.5f : p[2]};
            const float su = sample_ip<IP>(u.v[a], q[0], q[1], q[2], 1, bc);
            const float st = sample_ip<IP>(ut.v[a], q[0], q[1], q[2], 1, bc);
            const float delta = su - st;  // g2p.rs:43-44
            out[3 * i + a] = (u_old[3 * i + a] + delta) * alpha + su * (1.0f - alpha);
        }
    }
}

// ---------------------------------------------------------------------------
// Reference MAC P2G (fluid/p2g.rs:221-412).
// ---------------------------------------------------------------------------
__global__ void bucket_keys_kernel(const float *__restrict__ pos, int64_t n, int32_t nx, int32_t ny,
                                   uint32_t *__restrict__ keys, uint32_t *__restrict__ idx) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t x = trunc_index(pos[3 * i]), y = trunc_index(pos[3 * i + 1]), z = trunc_index(pos[3 * i + 2]);
        keys[i] = (x >> 4) + (y >> 4) * (uint32_t)nx + (z >> 4) * (uint32_t)nx * (uint32_t)ny;
        idx[i] = (uint32_t)i;
    }
}

__global__ void set_flag_kernel(int32_t *p, int32_t v) { *p = v; }

// bucket_keys_kernel, and *ungrouped set to 1 when some particle's block
// precedes the previous particle's in Morton order: otherwise (Morton-sorted
// input) each block's particles already form one run in input order, which is
// the order the stable bucket sort gives within a block.
__global__ void bucket_keys_grouped_kernel(const float *__restrict__ pos, int64_t n, int32_t nx, int32_t ny,
                                           uint32_t *__restrict__ keys, uint32_t *__restrict__ idx,
                                           int32_t *__restrict__ ungrouped) {
    auto mblock = [](uint32_t x, uint32_t y, uint32_t z) {
        return spread10(x >> 4) | (spread10(y >> 4) << 1) | (spread10(z >> 4) << 2);
    };
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t x = trunc_index(pos[3 * i]), y = trunc_index(pos[3 * i + 1]), z = trunc_index(pos[3 * i + 2]);
        keys[i] = (x >> 4) + (y >> 4) * (uint32_t)nx + (z >> 4) * (uint32_t)nx * (uint32_t)ny;
        idx[i] = (uint32_t)i;
        bool down = false;
        if (i > 0) {
            const uint32_t px = trunc_index(pos[3 * i - 3]), py = trunc_index(pos[3 * i - 2]),
                           pz = trunc_index(pos[3 * i - 1]);
            down = mblock(x, y, z) < mblock(px, py, pz);
        }
        if (__any(down) && (threadIdx.x & 63) == 0) *ungrouped = 1;  // one store per wave
    }
}

// Gathers (pos, vel, mass) rows into sorted order.
__global__ void gather_payload_kernel(const float *__restrict__ pos, const float *__restrict__ vel,
                                      const float *__restrict__ mass, const uint32_t *__restrict__ perm, int64_t n,
                                      float *__restrict__ spos, float *__restrict__ svel, float *__restrict__ smass) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = perm[i];
        spos[3 * i] = pos[3 * j];
        spos[3 * i + 1] = pos[3 * j + 1];
        spos[3 * i + 2] = pos[3 * j + 2];
        svel[3 * i] = vel[3 * j];
        svel[3 * i + 1] = vel[3 * j + 1];
        svel[3 * i + 2] = vel[3 * j + 2];
        if (smass) smass[i] = mass ? mass[j] : 1.0f;
    }
}

// A work item of the MAC splat: particles p0 .. p1 - 1 (sorted) of occupied
// block k; `shared` when the block's particles are split over several items.
struct SplatItem {
    int32_t k, p0, p1, shared;
};
constexpr int kSplatThreads = 1024;  // one group per CU holds the 144 KB window of r_p <= 2: 16 waves
constexpr int kSplatChunk = 4096;  // particles per item (4 per lane)

// Sum over the work-group (every thread receives it); s_red holds one slot per wave.
__device__ __forceinline__ float splat_wg_sum(float v, float *s_red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
    __syncthreads();
    float t = 0.0f;
    for (int w = 0; w < kSplatThreads / 64; ++w) t += s_red[w];
    __syncthreads();
    return t;
}

// One (direction a, colour) pass of the MAC splat (p2g.rs:261-333):
// workgroup i stages the particles of items[i] into its window and commits
// it (with global atomics when other items of the same colour pass share the
// block).  CH 0 splats the mass m w, CH 1 the momentum m u_a w, CH 2 both.
//
// The window accumulates in fixed point: with B >= the item's sum of
// |contribution| and B < 2^e, every contribution c is scaled by 2^S,
// S = 61 - e, so |window sum| 2^S < 2^61.  A scaled contribution of
// magnitude >= 2^24 is an integer in float, and its int64 LDS add is exact:
// those sums carry no rounding and do not depend on order.  The rare smaller
// ones (below ~2^-37 B: a particle at the edge of a cell's support) are added
// in float: into a float window (CH 0 / 1), or (CH 2, where two int64
// windows leave no room for it) into a list of kResCap entries applied at the
// commit; past kResCap they are rounded to the fixed-point grid (absolute
// error <= 2^-S-1 each).  The commit adds (double) acc 2^-S + the float part,
// rounded once.  Integer LDS atomics run ~3x faster than float ones here.
constexpr int kResCap = 2048;
constexpr int kSplatMaxCells = 8192;  // window cells of CH 2 (m <= 2: 20^3)

// |q| of an integral float, 2^24 <= |q| < 2^62, as a 64-bit integer:
// |q| = t 2^32 + lo with t = trunc(|q| 2^-32) and 0 <= lo < 2^32, both exact
// in float (below 2^32 lo is |q| itself; above, the ulp of |q| is >= 2^9, so
// lo needs at most 23 bits).
__device__ __forceinline__ unsigned long long integral_mag_u64(float q) {
    const float a = fabsf(q);
    const float t = truncf(a * 0x1p-32f);
    const float lo = __builtin_fmaf(t, -0x1p32f, a);
    return (unsigned long long)(uint32_t)t << 32 | (unsigned long long)(uint32_t)lo;
}

// AT is the face direction when it is a template constant (the one-pass
// CH 2 kernels), or -1 when it comes in `a_rt`.
template <int CH, int AT>
__global__ __launch_bounds__(kSplatThreads) void mac_splat_kernel(GridView gm, GridView gv, int a_rt,
                                                                  const SplatItem *__restrict__ items,
                                                                  const int32_t *__restrict__ pblocks,
                                                                  const float *__restrict__ spos,
                                                                  const float *__restrict__ svel,
                                                                  const float *__restrict__ smass, float r_p,
                                                                  float inv_r2, int m) {
    extern __shared__ __attribute__((aligned(8))) unsigned char smem_raw[];
    __shared__ float s_red[kSplatThreads / 64];
    __shared__ int s_nres;
    __shared__ int s_slot[2][27];  // mass / velocity slot of each committed region, -1 where not committed
    constexpr int NC = CH == 2 ? 2 : 1;  // channels of this pass
    const int a = AT >= 0 ? AT : a_rt;
    const int MSX = 16 + 2 * m, NS = MSX * MSX * MSX;
    unsigned long long *acc = reinterpret_cast<unsigned long long *>(smem_raw);  // NC windows
    float *res = reinterpret_cast<float *>(acc + NC * NS);                         // CH 0 / 1: float window
    uint2 *rlist = reinterpret_cast<uint2 *>(acc + NC * NS);                       // CH 2: (cell | channel << 31, value)
    for (int s = threadIdx.x; s < NC * NS; s += blockDim.x) acc[s] = 0ull;
    if (CH != 2)
        for (int s = threadIdx.x; s < NS; s += blockDim.x) res[s] = 0.0f;
    if (threadIdx.x == 0) s_nres = 0;
    const SplatItem item = items[blockIdx.x];
    const int bid = pblocks[item.k];
    const int bx = bid % gm.nx, by = (bid / gm.nx) % gm.ny, bz = bid / (gm.nx * gm.ny);
    const int b0x = bx * 16, b0y = by * 16, b0z = bz * 16;
    const int x0s = b0x - m, y0s = b0y - m, z0s = b0z - m, whi = 16 + m - 1;
    const float ofx = a == 0 ? 1.0f : 0.5f, ofy = a == 1 ? 1.0f : 0.5f, ofz = a == 2 ? 1.0f : 0.5f;
    if (threadIdx.x < 27) {
        // Commit regions (p2g.rs:342-411): window region (rx, ry, rz) is the
        // margin below, the block's own 16 cells or the margin above on each
        // axis; it is committed when it lies inside the domain and its block
        // holds both a mass and a velocity slot.
        const int t = threadIdx.x, rx = t % 3, ry = (t / 3) % 3, rz = t / 9;
        const int xlo = rx == 0 ? 0 : (rx == 1 ? m : m + 16), xlen = rx == 1 ? 16 : m;
        const int ylo = ry == 0 ? 0 : (ry == 1 ? m : m + 16), ylen = ry == 1 ? 16 : m;
        const int zlo = rz == 0 ? 0 : (rz == 1 ? m : m + 16), zlen = rz == 1 ? 16 : m;
        const int gx0 = x0s + xlo, gy0 = y0s + ylo, gz0 = z0s + zlo;
        int ms = -1, vs = -1;
        if (gx0 >= 0 && gy0 >= 0 && gz0 >= 0 && gx0 + xlen <= gm.sx && gy0 + ylen <= gm.sy && gz0 + zlen <= gm.sz) {
            const int64_t nb = (int64_t)(bx + rx - 1) + (int64_t)(by + ry - 1) * gm.nx +
                               (int64_t)(bz + rz - 1) * gm.nx * gm.ny;
            ms = gm.b2s[nb];
            vs = gv.b2s[nb];
            if (ms < 2 || vs < 2) ms = vs = -1;
        }
        s_slot[0][t] = ms;
        s_slot[1][t] = vs;
    }
    // Lane t takes a contiguous run of the item's particles.
    const int p_beg = item.p0, p_cnt = item.p1 - item.p0;
    const int per = (p_cnt + (int)blockDim.x - 1) / (int)blockDim.x;
    // channel k of this pass is the mass when CH is 0, or CH is 2 and k is 0
    auto is_mass = [](int k) { return CH == 2 ? k == 0 : CH == 0; };
    auto scale_of = [&](int pi, int k) { return is_mass(k) ? smass[pi] : smass[pi] * svel[3 * pi + a]; };
    // A particle whose mass or mass * velocity is not finite stays out of the
    // fixed-point window (its sum has no scale); it is added in float after
    // the commit, so only its own footprint turns non-finite.
    float bsum[NC];
    bool nonfin = false;
#pragma unroll
    for (int k = 0; k < NC; ++k) bsum[k] = 0.0f;
    for (int j = 0; j < per; ++j) {
        const int pi = p_beg + (int)threadIdx.x * per + j;
        if (pi < p_beg + p_cnt)
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                const float v = scale_of(pi, k);
                if (isfinite(v))
                    bsum[k] += fabsf(v);
                else
                    nonfin = true;
            }
    }
    // w <= 1; the FP32 sum of the bounds is within a factor 1 + 2^-12 of its
    // value, and e leaves a factor 2 spare.
    float up[NC];
    int ex[NC];
    bool any = false;
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        const float bound = splat_wg_sum(bsum[k], s_red);
        any = any || bound > 0.0f;
        frexpf(bound > 0.0f && isfinite(bound) ? bound : 1.0f, &ex[k]);  // bound < 2^e
        // S = 61 - e, at most 126: 2^S stays a float, and a contribution
        // reaching 2^24 after scaling comes from a normal float
        ex[k] = max(ex[k], 61 - 126);
        up[k] = ldexpf(1.0f, 61 - ex[k]);  // 2^S
    }
    const bool item_nonfin = splat_wg_sum(nonfin ? 1.0f : 0.0f, s_red) > 0.0f;  // uniform per group
    // S >= 0 on every channel: a finite nonzero scale times 2^S is nonzero,
    // so a zero scaled contribution is a zero contribution
    bool unit_scales = true;
#pragma unroll
    for (int k = 0; k < NC; ++k) unit_scales = unit_scales && ex[k] <= 61;
    // Nothing but zero contributions: the item adds nothing.  Uniform per group.
    if (!any && !item_nonfin) return;
    constexpr float kExact = 16777216.0f;  // 2^24
    auto footprint = [&](int pi, int &ix1, int &ix2, int &iy1, int &iy2, int &iz1, int &iz2) {
        const float px = spos[3 * pi], py = spos[3 * pi + 1], pz = spos[3 * pi + 2];
        ix1 = min(max((int)ceilf(px - r_p - ofx), b0x - m), b0x + whi);
        ix2 = min(max((int)floorf(px + r_p - ofx), b0x - m), b0x + whi);
        iy1 = min(max((int)ceilf(py - r_p - ofy), b0y - m), b0y + whi);
        iy2 = min(max((int)floorf(py + r_p - ofy), b0y - m), b0y + whi);
        iz1 = min(max((int)ceilf(pz - r_p - ofz), b0z - m), b0z + whi);
        iz2 = min(max((int)floorf(pz + r_p - ofz), b0z - m), b0z + whi);
    };
    // Eq. 6 weight of face cell (ix, iy, iz) for the particle at p0 = p - 1/2.
    auto weight = [&](float px0, float py0, float pz0, int ix, int iy, int iz) {
        const float dz = (float)iz - pz0, dy = (float)iy - py0, dx = (float)ix - px0;
        const float dist_sq = dx * dx + (dz * dz + dy * dy);
        const float d_a = a == 0 ? dx : (a == 1 ? dy : dz);
        const float df2 = dist_sq + d_a + 0.25f;
        const float w = fmaxf(1.0f - df2 * inv_r2, 0.0f);
        return w * w * w;
    };
    // One contribution w * sc[k] of a particle whose scales are finite (sf =
    // sc, sq = sc 2^S; both 0 for a non-finite channel).  q = w sq equals
    // (w sc) 2^S bit for bit whenever |q| >= 2^24 (2^S is a power of two and
    // w sc is then a normal float), so the window receives the same integers
    // as from scaling the float contribution.
    auto contrib = [&](int cell, float w, const float *sf, const float *sq) {
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            const float q = w * sq[k];
            if (fabsf(q) >= kExact) {
                const unsigned long long u = integral_mag_u64(q);
                if (q > 0.0f)
                    atomicAdd(&acc[k * NS + cell], u);
                else
                    __hip_atomic_fetch_sub(&acc[k * NS + cell], u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            } else {
                const float c = w * sf[k];
                if (c == 0.0f) continue;  // +-0 keeps the value
                if (CH != 2) {
                    atomicAdd(&res[cell], c);
                } else {
                    const int slot = atomicAdd(&s_nres, 1);
                    if (slot < kResCap)
                        rlist[slot] = make_uint2((uint32_t)cell | (uint32_t)k << 31, __float_as_uint(c));
                    else
                        atomicAdd(&acc[k * NS + cell], (unsigned long long)llrintf(c * up[k]));
                }
            }
        }
    };
    for (int j = 0; j < per; ++j) {
        const int pi = p_beg + (int)threadIdx.x * per + j;
        if (pi >= p_beg + p_cnt) break;
        float sf[NC], sq[NC];
        unsigned long long sgn[NC];  // all ones where the channel's contributions are negative
        bool live = false;
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            const float v = scale_of(pi, k);
            const bool fin = isfinite(v);  // non-finite: added after the commit
            sf[k] = fin ? v : 0.0f;
            sq[k] = fin ? v * up[k] : 0.0f;
            sgn[k] = sq[k] < 0.0f ? ~0ull : 0ull;
            live = live || (fin && v != 0.0f);
        }
        if (!live) continue;
        const float px0 = spos[3 * pi] - 0.5f, py0 = spos[3 * pi + 1] - 0.5f, pz0 = spos[3 * pi + 2] - 0.5f;
        int ix1, ix2, iy1, iy2, iz1, iz2;
        footprint(pi, ix1, ix2, iy1, iy2, iz1, iz2);
        if (AT >= 0 && unit_scales && ix2 - ix1 <= 3 && iy2 - iy1 <= 3 && iz2 - iz1 <= 3) {
            // r_p <= 2: the footprint is at most 4 cells wide on an axis
            // unless p - of is an integer there (5 wide: the loop below).
            // The 4^3 box is unrolled; offsets past the footprint get an
            // infinite squared distance, hence weight 0 and no LDS access.
            float cx[4], cx2[4], cy[4], cy2[4], cz[4], cz2[4];
#pragma unroll
            for (int o = 0; o < 4; ++o) {
                cx[o] = (float)(ix1 + o) - px0;
                cy[o] = (float)(iy1 + o) - py0;
                cz[o] = (float)(iz1 + o) - pz0;
                cx2[o] = ix1 + o <= ix2 ? cx[o] * cx[o] : __builtin_inff();
                cy2[o] = iy1 + o <= iy2 ? cy[o] * cy[o] : __builtin_inff();
                cz2[o] = iz1 + o <= iz2 ? cz[o] * cz[o] : __builtin_inff();
                // computed once per particle: the empty statements keep the
                // compiler from recomputing them at each of the 64 cells
                asm volatile("" : "+v"(cx[o]), "+v"(cx2[o]));
            }
            const int base = ((iz1 - z0s) * MSX + (iy1 - y0s)) * MSX + ix1 - x0s;
            // A cell can hold a contribution 0 < |q| < 2^24 only if its weight
            // is below wthr: |q| = fl(w |sq|) < 2^24 needs w |sq| < 2^24, and
            // wthr >= 2^24 / |sq| on every channel with sq != 0 (the quotient
            // rounded, then raised by 2^-21).  Cells with 0 < w < wthr take the
            // vote, a superset of those that need it.
            float wthr = 0.0f;
#pragma unroll
            for (int k = 0; k < NC; ++k)
                if (sq[k] != 0.0f) wthr = fmaxf(wthr, (kExact / fabsf(sq[k])) * (1.0f + 0x1p-21f));
#pragma unroll
            for (int oz = 0; oz < 4; ++oz)
#pragma unroll
                for (int oy = 0; oy < 4; ++oy) {
                    const float dyz = cz2[oz] + cy2[oy];
                    const int row = base + (oz * MSX + oy) * MSX;
#pragma unroll
                    for (int ox = 0; ox < 4; ++ox) {
                        const float dist_sq = cx2[ox] + dyz;
                        const float d_a = AT == 0 ? cx[ox] : (AT == 1 ? cy[oy] : cz[oz]);
                        const float df2 = dist_sq + d_a + 0.25f;
                        float w = fmaxf(1.0f - df2 * inv_r2, 0.0f);
                        w = w * w * w;
                        // Common case, no branch: every lane's scaled
                        // contribution is 0 or at least 2^24 in magnitude,
                        // so its magnitude converts exactly and the signed
                        // add (zero included) goes straight to the window.
                        if (__any((w > 0.0f) & (w < wthr))) {  // no short circuit
                            contrib(row + ox, w, sf, sq);
                        } else {
#pragma unroll
                            for (int k = 0; k < NC; ++k)
                                atomicAdd(&acc[k * NS + row + ox], (integral_mag_u64(w * sq[k]) ^ sgn[k]) - sgn[k]);
                        }
                    }
                }
        } else {
            for (int iz = iz1; iz <= iz2; ++iz)
                for (int iy = iy1; iy <= iy2; ++iy)
                    for (int ix = ix1; ix <= ix2; ++ix)
                        contrib(((iz - z0s) * MSX + (iy - y0s)) * MSX + ix - x0s, weight(px0, py0, pz0, ix, iy, iz),
                                sf, sq);
        }
    }
    __syncthreads();
    // Window values in float: CH 0 / 1 into the residue window; CH 2 into
    // the first half of each int64 window's bytes (read all, barrier, write),
    // then the listed small terms on top.
    float *fv[NC];
    if (CH != 2) {
        const double down = ldexp(1.0, ex[0] - 61);
        for (int s = threadIdx.x; s < NS; s += blockDim.x) {
            const long long ia = (long long)acc[s];
            res[s] = (float)((double)ia * down + (double)res[s]);
        }
        fv[0] = res;
    } else {
        float v[NC][kSplatMaxCells / kSplatThreads];
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            const double down = ldexp(1.0, ex[k] - 61);
#pragma unroll
            for (int i = 0; i < kSplatMaxCells / kSplatThreads; ++i) {
                const int s = threadIdx.x + i * kSplatThreads;
                v[k][i] = s < NS ? (float)((double)(long long)acc[k * NS + s] * down) : 0.0f;
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            fv[k] = reinterpret_cast<float *>(acc + k * NS);
#pragma unroll
            for (int i = 0; i < kSplatMaxCells / kSplatThreads; ++i) {
                const int s = threadIdx.x + i * kSplatThreads;
                if (s < NS) fv[k][s] = v[k][i];
            }
        }
        __syncthreads();
        const int nl = min(s_nres, kResCap);
        for (int i = threadIdx.x; i < nl; i += blockDim.x) {
            const uint2 e = rlist[i];
            atomicAdd(&fv[e.x >> 31][e.x & 0x7fffffffu], __uint_as_float(e.y));
        }
    }
    __syncthreads();
    // Commit (p2g.rs:342-411): staging cell -> its region -> the region's
    // block (the region table above).  Window cell s = (lz MSX + ly) MSX + lx
    // is split in float: (s + 1/2) / n is at least 1/(2n) away from an
    // integer, far beyond its rounding.
    const float inv_m1 = 1.0f / (float)MSX, inv_m2 = 1.0f / (float)(MSX * MSX);
    auto region = [&](int l) { return l < m ? 0 : (l < m + 16 ? 1 : 2); };
    // The target of staging cell s for channel k, or null where its region is not committed.
    auto target = [&](int s, int k) -> float * {
        const int lz = (int)(((float)s + 0.5f) * inv_m2), r = s - lz * MSX * MSX;
        const int ly = (int)(((float)r + 0.5f) * inv_m1), lx = r - ly * MSX;
        const int slot = s_slot[is_mass(k) ? 0 : 1][region(lx) + 3 * region(ly) + 9 * region(lz)];
        if (slot < 0) return nullptr;
        const int vid = ((x0s + lx) & 15) | (((y0s + ly) & 15) << 4) | (((z0s + lz) & 15) << 8);
        return (is_mass(k) ? gm.slab : gv.slab) + (int64_t)slot * 4096 + vid;
    };
    // Four window cells per lane at a time, their reads issued before their writes.
    for (int s0 = threadIdx.x; s0 < NS; s0 += 4 * kSplatThreads) {
        float *pt[4][NC];
        float val[4][NC];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int s = s0 + i * kSplatThreads;
            const bool in = s < NS;
            const int sc = in ? s : 0;
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                val[i][k] = in ? fv[k][sc] : 0.0f;
                pt[i][k] = val[i][k] != 0.0f ? target(sc, k) : nullptr;
            }
        }
        if (item.shared) {  // blocks of one colour never share a window: only the block's own items race
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int k = 0; k < NC; ++k)
                    if (pt[i][k]) unsafeAtomicAdd(pt[i][k], val[i][k]);
        } else {
            float old[4][NC];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int k = 0; k < NC; ++k) old[i][k] = pt[i][k] ? *pt[i][k] : 0.0f;
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int k = 0; k < NC; ++k)
                    if (pt[i][k]) *pt[i][k] = old[i][k] + val[i][k];
        }
    }
    if (!item_nonfin) return;
    // Particles with a non-finite mass or momentum scale: their float
    // contributions go straight to the committed cells (NaN / Inf propagate
    // as in a float sum; finite cells outside their footprints keep the exact
    // result above).
    __syncthreads();
    for (int j = 0; j < per; ++j) {
        const int pi = p_beg + (int)threadIdx.x * per + j;
        if (pi >= p_beg + p_cnt) break;
        float sc[NC];
        bool bad = false;
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            sc[k] = scale_of(pi, k);
            bad = bad || !isfinite(sc[k]);
        }
        if (!bad) continue;
        const float px0 = spos[3 * pi] - 0.5f, py0 = spos[3 * pi + 1] - 0.5f, pz0 = spos[3 * pi + 2] - 0.5f;
        int ix1, ix2, iy1, iy2, iz1, iz2;
        footprint(pi, ix1, ix2, iy1, iy2, iz1, iz2);
        for (int iz = iz1; iz <= iz2; ++iz)
            for (int iy = iy1; iy <= iy2; ++iy)
                for (int ix = ix1; ix <= ix2; ++ix) {
                    const float w = weight(px0, py0, pz0, ix, iy, iz);
                    const int cell = ((iz - z0s) * MSX + (iy - y0s)) * MSX + ix - x0s;
#pragma unroll
                    for (int k = 0; k < NC; ++k) {
                        if (isfinite(sc[k])) continue;
                        float *po = target(cell, k);
                        if (po) atomicAdd(po, w * sc[k]);
                    }
                }
    }
}

// MacGrid::normalize_velocity (mac.rs:106-132): u = P/M, 0 where M <= 0.
__global__ void normalize_kernel(GridView vg, GridView mg, const int32_t *__restrict__ active) {
    const int64_t bid = active[blockIdx.x];
    const int vs = vg.b2s[bid], ms = mg.b2s[bid];
    if (vs < 2) return;
    for (int j = threadIdx.x; j < 4096; j += blockDim.x) {
        const float mm = mg.slab[(int64_t)ms * 4096 + j];
        float &v = vg.slab[(int64_t)vs * 4096 + j];
        v = mm > 0.0f ? v / mm : 0.0f;
    }
}

// ---------------------------------------------------------------------------
// B-spline transfers.
// ---------------------------------------------------------------------------
template <int ORDER>
__device__ __forceinline__ float bspline_w(float p, int ci) {
    const float x = (p - 0.5f) - (float)ci;
    const float ax = fabsf(x);
    if (ORDER == 1) return ax < 1.0f ? 1.0f - ax : 0.0f;
    if (ax < 0.5f) return 0.75f - x * x;
    if (ax < 1.5f) {
        const float t = 1.5f - ax;
        return 0.5f * t * t;
    }
    return 0.0f;
}

// Sorted particle records for the B-spline P2G, pm = (x, y, z, m) and
// pu = (u, v, w, 0), gathered from the caller-order records rec.
__global__ void gather_records_kernel(const float4 *__restrict__ rec, const uint32_t *__restrict__ perm, int64_t n,
                                      float4 *__restrict__ pm, float4 *__restrict__ pu) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = perm[i];
        pm[i] = rec[2 * j];
        pu[i] = rec[2 * j + 1];
    }
}

// Voxel ranges of the sorted particles: vbeg/vend indexed by slot*4096+vid.
__global__ void voxel_ranges_kernel(const uint32_t *__restrict__ keys, const float4 *__restrict__ pm, int64_t n,
                                    GridView g, uint32_t *__restrict__ vbeg, uint32_t *__restrict__ vend) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const bool head = i == 0 || keys[i - 1] != keys[i];
        const bool tail = i == n - 1 || keys[i + 1] != keys[i];
        if (!head && !tail) continue;
        const float4 p = pm[i];
        const uint32_t x = trunc_index(p.x), y = trunc_index(p.y), z = trunc_index(p.z);
        if (x >= (uint32_t)g.sx || y >= (uint32_t)g.sy || z >= (uint32_t)g.sz) continue;
        const int slot = g.b2s[(x >> 4) + (int64_t)(y >> 4) * g.nx + (int64_t)(z >> 4) * g.nx * g.ny];
        if (slot < 2) continue;
        const int64_t c = (int64_t)slot * 4096 + ((x & 15) | ((y & 15) << 4) | ((z & 15) << 8));
        if (head) vbeg[c] = (uint32_t)i;
        if (tail) vend[c] = (uint32_t)(i + 1);
    }
}

struct Acc4 {
    float m, u0, u1, u2;
};

template <int ORDER>
__device__ __forceinline__ void acc_add(Acc4 &a, float wxy, float pz, int cz, float m, const float4 &u) {
    const float wm = (wxy * bspline_w<ORDER>(pz, cz)) * m;
    a.m = a.m + wm;
    a.u0 = a.u0 + wm * u.x;
    a.u1 = a.u1 + wm * u.y;
    a.u2 = a.u2 + wm * u.z;
}

// Workgroup per active block; lane (lx, ly) owns the cell column (lx, ly, *).
// A cell sums its 27 neighbour voxels, z then y then x from -1 to 1, and each
// voxel's particles in sorted order -- the oracle's order.  The lane walks the
// 18 voxel planes bz*16-1 .. bz*16+16 once: plane v feeds cells v+1 (as their
// dz = -1 plane), v (dz = 0) and v-1 (dz = +1), so each particle is read once
// per column and every cell still receives its planes in increasing z.
template <int ORDER>
__global__ __launch_bounds__(256) void p2g_bspline_kernel(GridView mg, GridView v0, GridView v1, GridView v2,
                                                          const int32_t *__restrict__ active,
                                                          const uint32_t *__restrict__ vbeg,
                                                          const uint32_t *__restrict__ vend,
                                                          const float4 *__restrict__ pm,
                                                          const float4 *__restrict__ pu) {
    __shared__ int s_nb[27];
    // voxel ranges of the block's 18 x 18 plane window (the block's columns
    // and a 1-voxel ring), double-buffered: plane v + 1 loads during plane v
    __shared__ uint2 s_pl[2][18 * 18];
    const int64_t bid = active[blockIdx.x];
    const int bx = (int)(bid % mg.nx), by = (int)((bid / mg.nx) % mg.ny), bz = (int)(bid / ((int64_t)mg.nx * mg.ny));
    if (threadIdx.x < 27) {
        const int t = threadIdx.x;
        const int nx = bx + t % 3 - 1, ny = by + (t / 3) % 3 - 1, nz = bz + t / 9 - 1;
        s_nb[t] = (nx >= 0 && nx < mg.nx && ny >= 0 && ny < mg.ny && nz >= 0 && nz < mg.nz)
                      ? mg.b2s[nx + (int64_t)ny * mg.nx + (int64_t)nz * mg.nx * mg.ny]
                      : 0;
    }
    __syncthreads();
    const int ms = s_nb[13];
    const int s0 = v0.b2s[bid], s1 = v1.b2s[bid], s2 = v2.b2s[bid];
    if (ms < 2) return;
    const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
    const int cx = bx * 16 + lx, cy = by * 16 + ly, cz0 = bz * 16;
    Acc4 lo = {0, 0, 0, 0}, mid = {0, 0, 0, 0}, hi = {0, 0, 0, 0};  // cells lv-1, lv, lv+1
    auto store = [&](int lz, const Acc4 &a) {
        const int vid = lx | (ly << 4) | (lz << 8);
        mg.slab[(int64_t)ms * 4096 + vid] = a.m;
        const bool pos_m = a.m > 0.0f;
        v0.slab[(int64_t)s0 * 4096 + vid] = pos_m ? a.u0 / a.m : 0.0f;
        v1.slab[(int64_t)s1 * 4096 + vid] = pos_m ? a.u1 / a.m : 0.0f;
        v2.slab[(int64_t)s2 * 4096 + vid] = pos_m ? a.u2 / a.m : 0.0f;
    };
    // Ranges of window voxels idx and idx + 256 (idx < 324) of plane vz; an
    // outside or absent voxel is an empty range.
    auto plane_ranges = [&](int vz, uint2 *r) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int idx = threadIdx.x + 256 * h;
            r[h] = make_uint2(0u, 0u);
            if (idx >= 18 * 18 || vz < 0 || vz >= mg.sz) continue;
            const int vy = by * 16 - 1 + idx / 18, vx = bx * 16 - 1 + idx % 18;
            if (vy < 0 || vy >= mg.sy || vx < 0 || vx >= mg.sx) continue;
            const int slot = s_nb[((vx >> 4) - bx + 1) + 3 * ((vy >> 4) - by + 1) + 9 * ((vz >> 4) - bz + 1)];
            if (slot < 2) continue;
            const int64_t c = (int64_t)slot * 4096 + ((vx & 15) | ((vy & 15) << 4) | ((vz & 15) << 8));
            r[h] = make_uint2(vbeg[c], vend[c]);
        }
    };
    auto put_plane = [&](int buf, const uint2 *r) {
        s_pl[buf][threadIdx.x] = r[0];
        if (threadIdx.x + 256 < 18 * 18) s_pl[buf][threadIdx.x + 256] = r[1];
    };
    uint2 nxt[2];
    plane_ranges(cz0 - 1, nxt);
    put_plane(0, nxt);
    __syncthreads();
    for (int lv = -1; lv <= 16; ++lv) {
        const int vz = cz0 + lv, cur = (lv + 1) & 1;
        if (lv < 16) plane_ranges(vz + 1, nxt);  // in flight while this plane's particles are added
        if (vz >= 0 && vz < mg.sz) {
            // This lane's 9 voxel ranges, (dy, dx) in order, are window
            // voxels (ly + dy + 1, lx + dx + 1).  One pass over their
            // concatenation: a wave steps as often as its fullest lane has
            // particles, not once per voxel per the fullest lane of that
            // voxel.  Per cell, the particles still arrive voxel by voxel in
            // (dy, dx) order, sorted within.
            const uint2 *pl = s_pl[cur] + ly * 18 + lx;
            auto rng = [&](int t) { return pl[(t / 3) * 18 + t % 3]; };
            int t = 0;
            uint2 rt = rng(0);
            uint32_t j = rt.x, e = rt.y;
            while (j >= e && ++t < 9) {
                rt = rng(t);
                j = rt.x;
                e = rt.y;
            }
            float4 pn = make_float4(0.f, 0.f, 0.f, 0.f), un = pn;
            if (t < 9) {
                pn = pm[j];
                un = pu[j];
            }
            while (t < 9) {
                const float4 p = pn, u = un;
                ++j;
                while (j >= e && ++t < 9) {
                    rt = rng(t);
                    j = rt.x;
                    e = rt.y;
                }
                if (t < 9) {  // the next particle's records are in flight while this one is added
                    pn = pm[j];
                    un = pu[j];
                }
                const float wxy = bspline_w<ORDER>(p.x, cx) * bspline_w<ORDER>(p.y, cy);
                if (lv >= 1) acc_add<ORDER>(lo, wxy, p.z, cz0 + lv - 1, p.w, u);
                if (lv >= 0 && lv <= 15) acc_add<ORDER>(mid, wxy, p.z, cz0 + lv, p.w, u);
                if (lv <= 14) acc_add<ORDER>(hi, wxy, p.z, cz0 + lv + 1, p.w, u);
            }
        }
        if (lv < 16) put_plane(cur ^ 1, nxt);  // buffer cur ^ 1 was last read before the previous barrier
        __syncthreads();
        if (lv >= 1) store(lv - 1, lo);  // cell lv-1 has its planes lv-2, lv-1, lv
        lo = mid;
        mid = hi;
        hi = Acc4{0, 0, 0, 0};
    }
}

// G2P work-groups: 256 consecutive Morton-sorted particles each.  The grid
// cells their supports reach (the chunk's box of floor(p) -1 .. +1, clipped
// to the domain) are staged in LDS when the box holds at most kG2pWindow
// cells; a chunk with a larger box (it straddles a coarse Morton boundary)
// reads the slabs directly.  Both paths take the same values in the same
// order, so the result does not depend on which one runs.
constexpr int kG2pThreads = 256;
constexpr int kG2pWindow = 3072;
constexpr int kG2pWindowBlocks = 64;  // 16^3 blocks a staged box may touch

// floor(p) as a cell index, held to [-2, s + 1]: beyond that every support
// cell floor(p) - 1 .. floor(p) + 1 lies outside the domain either way, and
// the float -> int conversion stays in range (NaN maps to -2: no cells).
__device__ __forceinline__ int floor_cell(float p, int s) {
    return (int)fminf(fmaxf(floorf(p), -2.0f), (float)(s + 1));
}

// Cell range [lo, hi] of one axis of a particle's support, clipped to the
// domain (empty when lo > hi).
__device__ __forceinline__ void support_range(int f, int s, int &lo, int &hi) {
    lo = max(f - 1, 0);
    hi = min(f + 1, s - 1);
}

// One particle's sum over its 27 support cells (z, then y, then x), the
// oracle's order.  STAGED reads the cells from the LDS window `win` (box
// origin x0, y0, z0, extents ex, ey), else straight from the slabs.
template <int ORDER, bool STAGED>
__device__ __forceinline__ void g2p_sum(const GridView &v0, const GridView &v1, const GridView &v2, float px,
                                        float py, float pz, int fx, int fy, int fz, const float (*win)[kG2pWindow],
                                        int x0, int y0, int z0, int ex, int ey, float &a0, float &a1, float &a2) {
    a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
    for (int dz = -1; dz <= 1; ++dz) {
        const int cz = fz + dz;
        if (cz < 0 || cz >= v0.sz) continue;
        const float wz = bspline_w<ORDER>(pz, cz);
        for (int dy = -1; dy <= 1; ++dy) {
            const int cy = fy + dy;
            if (cy < 0 || cy >= v0.sy) continue;
            const float wy = bspline_w<ORDER>(py, cy);
            for (int dx = -1; dx <= 1; ++dx) {
                const int cx = fx + dx;
                if (cx < 0 || cx >= v0.sx) continue;
                const float w = (bspline_w<ORDER>(px, cx) * wy) * wz;
                float s0, s1, s2;
                if constexpr (STAGED) {  // every in-domain support cell lies in the box
                    const int c = ((cz - z0) * ey + (cy - y0)) * ex + (cx - x0);
                    s0 = win[0][c];
                    s1 = win[1][c];
                    s2 = win[2][c];
                } else {
                    const int64_t bid =
                        (cx >> 4) + (int64_t)(cy >> 4) * v0.nx + (int64_t)(cz >> 4) * v0.nx * v0.ny;
                    const int vid = (cx & 15) | ((cy & 15) << 4) | ((cz & 15) << 8);
                    s0 = v0.slab[(int64_t)v0.b2s[bid] * 4096 + vid];
                    s1 = v1.slab[(int64_t)v1.b2s[bid] * 4096 + vid];
                    s2 = v2.slab[(int64_t)v2.b2s[bid] * 4096 + vid];
                }
                a0 = a0 + w * s0;
                a1 = a1 + w * s1;
                a2 = a2 + w * s2;
            }
        }
    }
}

template <int ORDER>
__global__ __launch_bounds__(kG2pThreads) void g2p_bspline_kernel(GridView v0, GridView v1, GridView v2,
                                                                  const float *__restrict__ pos, int64_t n,
                                                                  const uint32_t *__restrict__ visit,
                                                                  float *__restrict__ out) {
    __shared__ float win[3][kG2pWindow];
    __shared__ int s_box[6];                    // min x, y, z; max x, y, z of the chunk's support cells
    __shared__ int s_slot[3][kG2pWindowBlocks];  // slab slot of each block the box touches, per component
    // Consecutive chunks (neighbouring cells) land on one XCD and share its L2.
    const int64_t k = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * kG2pThreads + threadIdx.x;
    if (threadIdx.x < 3) s_box[threadIdx.x] = INT_MAX;
    else if (threadIdx.x < 6) s_box[threadIdx.x] = INT_MIN;
    int64_t i = -1;
    float px = 0.0f, py = 0.0f, pz = 0.0f;
    int fx = 0, fy = 0, fz = 0;
    // support box of this lane, inverted (no cells) unless it reaches the domain
    int lx = INT_MAX, ly = INT_MAX, lz = INT_MAX, hx = INT_MIN, hy = INT_MIN, hz = INT_MIN;
    if (k < n) {
        i = visit[k];
        px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
        fx = floor_cell(px, v0.sx), fy = floor_cell(py, v0.sy), fz = floor_cell(pz, v0.sz);
        int ax, bx, ay, by, az, bz;
        support_range(fx, v0.sx, ax, bx);
        support_range(fy, v0.sy, ay, by);
        support_range(fz, v0.sz, az, bz);
        if (ax <= bx && ay <= by && az <= bz) lx = ax, hx = bx, ly = ay, hy = by, lz = az, hz = bz;
    }
    // wave-level min / max first: one LDS atomic per wave and bound
    for (int o = 32; o > 0; o >>= 1) {
        lx = min(lx, __shfl_xor(lx, o)), ly = min(ly, __shfl_xor(ly, o)), lz = min(lz, __shfl_xor(lz, o));
        hx = max(hx, __shfl_xor(hx, o)), hy = max(hy, __shfl_xor(hy, o)), hz = max(hz, __shfl_xor(hz, o));
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0 && lx <= hx) {
        atomicMin(&s_box[0], lx);
        atomicMin(&s_box[1], ly);
        atomicMin(&s_box[2], lz);
        atomicMax(&s_box[3], hx);
        atomicMax(&s_box[4], hy);
        atomicMax(&s_box[5], hz);
    }
    __syncthreads();
    const int x0 = s_box[0], y0 = s_box[1], z0 = s_box[2];
    // box extents; none of the chunk's lanes reaching the domain leaves the box inverted (ex <= 0)
    const int64_t ex = (int64_t)s_box[3] - x0 + 1, ey = (int64_t)s_box[4] - y0 + 1, ez = (int64_t)s_box[5] - z0 + 1;
    const int bx0 = x0 >> 4, by0 = y0 >> 4, bz0 = z0 >> 4;
    const int64_t nbx = ex > 0 ? ((s_box[3] >> 4) - bx0 + 1) : 0, nby = ex > 0 ? ((s_box[4] >> 4) - by0 + 1) : 0,
                  nbz = ex > 0 ? ((s_box[5] >> 4) - bz0 + 1) : 0;
    float a0, a1, a2;
    // uniform over the work-group
    if (ex > 0 && ex * ey * ez <= kG2pWindow && nbx * nby * nbz <= kG2pWindowBlocks) {
        const int nbxy = (int)(nbx * nby), nb = (int)(nbxy * nbz);
        if (threadIdx.x < nb) {
            const int t = threadIdx.x;
            const int64_t bid = (bx0 + t % (int)nbx) + (int64_t)(by0 + (t % nbxy) / (int)nbx) * v0.nx +
                                (int64_t)(bz0 + t / nbxy) * v0.nx * v0.ny;
            s_slot[0][t] = v0.b2s[bid];
            s_slot[1][t] = v1.b2s[bid];
            s_slot[2][t] = v2.b2s[bid];
        }
        __syncthreads();
        // Each lane's cells are addressed first and loaded back to back.
        constexpr int ITEMS = kG2pWindow / kG2pThreads;
        const int exy = (int)(ex * ey), nw = (int)(exy * ez);
        int64_t off[3][ITEMS];
#pragma unroll
        for (int q = 0; q < ITEMS; ++q) {
            const int c = threadIdx.x + q * kG2pThreads;
            off[0][q] = off[1][q] = off[2][q] = -1;
            if (c < nw) {
                const int cz = z0 + c / exy, cy = y0 + (c % exy) / (int)ex, cx = x0 + c % (int)ex;
                const int b = ((cx >> 4) - bx0) + (int)nbx * ((cy >> 4) - by0) + nbxy * ((cz >> 4) - bz0);
                const int vid = (cx & 15) | ((cy & 15) << 4) | ((cz & 15) << 8);
                off[0][q] = (int64_t)s_slot[0][b] * 4096 + vid;
                off[1][q] = (int64_t)s_slot[1][b] * 4096 + vid;
                off[2][q] = (int64_t)s_slot[2][b] * 4096 + vid;
            }
        }
        float val[3][ITEMS];
#pragma unroll
        for (int q = 0; q < ITEMS; ++q) {
            val[0][q] = off[0][q] >= 0 ? v0.slab[off[0][q]] : 0.0f;
            val[1][q] = off[1][q] >= 0 ? v1.slab[off[1][q]] : 0.0f;
            val[2][q] = off[2][q] >= 0 ? v2.slab[off[2][q]] : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < ITEMS; ++q) {
            const int c = threadIdx.x + q * kG2pThreads;
            if (c < nw) win[0][c] = val[0][q], win[1][c] = val[1][q], win[2][c] = val[2][q];
        }
        __syncthreads();
        if (k >= n) return;
        g2p_sum<ORDER, true>(v0, v1, v2, px, py, pz, fx, fy, fz, win, x0, y0, z0, (int)ex, (int)ey, a0, a1, a2);
    } else {
        if (k >= n) return;
        g2p_sum<ORDER, false>(v0, v1, v2, px, py, pz, fx, fy, fz, win, 0, 0, 0, 0, 0, a0, a1, a2);
    }
    out[3 * i] = a0;
    out[3 * i + 1] = a1;
    out[3 * i + 2] = a2;
}

// Ensures `bids` (host) on every grid and zero-fills them (MacGrid::zero_blocks,
// mac.rs:95-102).
int zero_blocks(msbg_grid *const *grids, int ng, const std::vector<int64_t> &bids) {
    for (int k = 0; k < ng; ++k) {
        int rc = msbg_grid_ensure_blocks(grids[k], bids.data(), (int64_t)bids.size());
        if (rc) return rc;
        rc = msbg_grid_fill_blocks(grids[k], bids.data(), (int64_t)bids.size(), 0.0f);
        if (rc) return rc;
    }
    return MSBG_OK;
}

int check_same_dims(msbg_grid *const *grids, int ng) {
    for (int k = 0; k < ng; ++k) {
        MSBG_CHECK(grids[k], MSBG_EINVAL, "transfer: null grid");
        MSBG_CHECK(grids[k]->L == 4 && grids[k]->sx == grids[0]->sx && grids[k]->sy == grids[0]->sy &&
                       grids[k]->sz == grids[0]->sz,
                   MSBG_EINVAL, "transfer: all channels must be 16^3-block grids over one domain");
    }
    return MSBG_OK;
}

// Active blocks of the particles (device list), downloaded.
int active_list(const msbg_grid *g, const float *d_pos, int64_t n, float r_scan, std::vector<int64_t> &out,
                int32_t **d_list) {
    int rc;
    int32_t *d_act = (int32_t *)workspace(6, sizeof(int32_t) * std::max<int64_t>(g->n_blocks, 1), &rc);
    if (rc) return rc;
    int64_t k = 0;
    rc = msbg_active_blocks(g->sx, g->sy, g->sz, d_pos, n, r_scan, d_act, &k);
    if (rc) return rc;
    std::vector<int32_t> h((size_t)k);
    if (k) MSBG_HIP(hipMemcpy(h.data(), d_act, sizeof(int32_t) * k, hipMemcpyDeviceToHost));
    out.assign(h.begin(), h.end());
    *d_list = d_act;
    return MSBG_OK;
}

}  // namespace msbg

using namespace msbg;

extern "C" {

int msbg_sample(const msbg_grid *g, const float *d_pos, int64_t n, int interp, int align, int bc, float *d_out) {
    MSBG_CHECK(g && g->L == 4, MSBG_EINVAL, "sample: 16^3-block grid required");
    MSBG_CHECK(interp == 0 || interp == 1, MSBG_EINVAL, "sample: interp is 0 (Linear) or 1 (CubicBSpline)");
    MSBG_CHECK((align == 0 || align == 1) && bc >= 0 && bc <= 2, MSBG_EINVAL, "sample: bad alignment / boundary");
    MSBG_CHECK(n >= 0 && (n == 0 || (d_pos && d_out)), MSBG_EINVAL, "sample: null argument");
    if (n == 0) return MSBG_OK;
    launch_sample(g->view(), d_pos, n, interp, align, bc, -1, d_out, 1);
    MSBG_HIP(hipGetLastError());
    MSBG_HIP(hipStreamSynchronize(stream()));
    return MSBG_OK;
}

int msbg_mac_sample(msbg_grid *const *vel, const float *d_pos, int64_t n, int interp, int bc, float *d_out) {
    MSBG_CHECK(vel, MSBG_EINVAL, "mac sample: null grids");
    int rc = check_same_dims(vel, 3);
    if (rc) return rc;
    MSBG_CHECK(interp == 0 || interp == 1, MSBG_EINVAL, "mac sample: interp is 0 or 1");
    MSBG_CHECK(bc >= 0 && bc <= 2, MSBG_EINVAL, "mac sample: bad boundary");
    MSBG_CHECK(n >= 0 && (n == 0 || (d_pos && d_out)), MSBG_EINVAL, "mac sample: null argument");
    if (n == 0) return MSBG_OK;
    for (int a = 0; a < 3; ++a)
        launch_sample(vel[a]->view(), d_pos, n, interp, 1, bc, a, d_out + a, 3);
    MSBG_HIP(hipGetLastError());
    MSBG_HIP(hipStreamSynchronize(stream()));
    return MSBG_OK;
}

int msbg_g2p_blend(const float *d_u_old, const float *d_u, const float *d_u_tilde, int64_t n, float alpha,
                   float *d_out) {
    MSBG_CHECK(n >= 0 && (n == 0 || (d_u_old && d_u && d_u_tilde && d_out)), MSBG_EINVAL, "g2p blend: null argument");
    if (n == 0) return MSBG_OK;
    hipLaunchKernelGGL(blend_kernel, dim3(grid_for(3 * n, 256)), dim3(256), 0, stream(), d_u_old, d_u, d_u_tilde, 3 * n,
                       alpha, d_out);
    MSBG_HIP(hipGetLastError());
    MSBG_HIP(hipStreamSynchronize(stream()));
    return MSBG_OK;
}

int msbg_g2p_flip(msbg_grid *const *vel, msbg_grid *const *vel_tilde, const float *d_pos, const float *d_u_old,
                  int64_t n, int interp, int bc, float alpha, float *d_out) {
    MSBG_CHECK(vel && vel_tilde, MSBG_EINVAL, "g2p flip: null grids");
    msbg_grid *all[6] = {vel[0], vel[1], vel[2], vel_tilde[0], vel_tilde[1], vel_tilde[2]};
    int rc = check_same_dims(all, 6);
    if (rc) return rc;
    MSBG_CHECK(interp == 0 || interp == 1, MSBG_EINVAL, "g2p flip: interp is 0 or 1");
    MSBG_CHECK(bc >= 0 && bc <= 2, MSBG_EINVAL, "g2p flip: bad boundary");
    MSBG_CHECK(n >= 0 && (n == 0 || (d_pos && d_u_old && d_out)), MSBG_EINVAL, "g2p flip: null argument");
    if (n == 0) return MSBG_OK;
    MacViews u, ut;
    for (int a = 0; a < 3; ++a) {
        u.v[a] = vel[a]->view();
        ut.v[a] = vel_tilde[a]->view();
    }
    if (interp == 0)
        hipLaunchKernelGGL(g2p_flip_kernel<0>, dim3(grid_for(n, 256)), dim3(256), 0, stream(), u, ut, d_pos, d_u_old,
                           n, bc, alpha, d_out);
    else
        hipLaunchKernelGGL(g2p_flip_kernel<1>, dim3(grid_for(n, 256)), dim3(256), 0, stream(), u, ut, d_pos, d_u_old,
                           n, bc, alpha, d_out);
    MSBG_HIP(hipGetLastError());
    MSBG_HIP(hipStreamSynchronize(stream()));
    return MSBG_OK;
}

int msbg_particles_to_grid_mac(msbg_grid *const *vel, msbg_grid *const *mass, const float *d_pos, const float *d_vel,
                               const float *d_mass, int64_t n, float r_p, int64_t *n_active) {
    MSBG_CHECK(vel && mass && n_active, MSBG_EINVAL, "particles_to_grid: null argument");
    msbg_grid *all[6] = {vel[0], vel[1], vel[2], mass[0], mass[1], mass[2]};
    int rc = check_same_dims(all, 6);
    if (rc) return rc;
    // splat precondition (p2g.rs:32-42, 174-181): 0 < r_p <= margin, staging fits LDS.
    MSBG_CHECK(r_p > 0.0f && r_p <= 3.0f, MSBG_EINVAL, "particles_to_grid: r_p must be in (0, 3]");
    MSBG_CHECK(n >= 0 && n < ((int64_t)1 << 31) && (n == 0 || (d_pos && d_vel && d_mass)), MSBG_EINVAL,
               "particles_to_grid: bad particle arrays");
    const msbg_grid *g = vel[0];
    // active blocks with the (r_p + 1)/bsx footprint (p2g.rs:104-107), zeroed
    std::vector<int64_t> act;
    int32_t *d_act = nullptr;
    rc = active_list(g, d_pos, n, r_p + 1.0f, act, &d_act);
    if (rc) return rc;
    rc = zero_blocks(all, 6, act);
    if (rc) return rc;
    *n_active = (int64_t)act.size();
    if (n == 0) return MSBG_OK;
    // bucket by block (p2g.rs:149-158).  Particles that arrive grouped by
    // block (Morton-sorted, as a FLIP step passes them) skip the sort and the
    // payload gather: the runs of the input are the buckets.
    uint32_t *keys = (uint32_t *)workspace(4, sizeof(uint32_t) * 2 * n + 2 * sizeof(int32_t), &rc);
    if (rc) return rc;
    uint32_t *perm = keys + n;
    int32_t *d_ungrouped = (int32_t *)(perm + n);
    hipLaunchKernelGGL(set_flag_kernel, dim3(1), dim3(1), 0, stream(), d_ungrouped, 0);
    hipLaunchKernelGGL(bucket_keys_grouped_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream(), d_pos, n,
                       (int32_t)g->nx, (int32_t)g->ny, keys, perm, d_ungrouped);
    MSBG_HIP(hipGetLastError());
    int32_t ungrouped = 1;
    MSBG_HIP(hipMemcpyAsync(&ungrouped, d_ungrouped, sizeof(int32_t), hipMemcpyDeviceToHost, stream()));
    MSBG_HIP(hipStreamSynchronize(stream()));
    float *pay = (float *)workspace(7, sizeof(float) * 7 * n + sizeof(int32_t) * (2 * n + 2), &rc);
    if (rc) return rc;
    const float *spos = d_pos, *svel = d_vel, *smass = d_mass;
    int32_t *pblocks = (int32_t *)(pay + 7 * n), *starts = pblocks + n;
    if (ungrouped) {
        int key_bits = 1;
        while (key_bits < 32 && ((int64_t)1 << key_bits) < g->n_blocks) ++key_bits;
        rc = radix_sort_pairs(keys, perm, n, key_bits);
        if (rc) return rc;
        hipLaunchKernelGGL(gather_payload_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream(), d_pos, d_vel, d_mass,
                           perm, n, pay, pay + 3 * n, pay + 6 * n);
        MSBG_HIP(hipGetLastError());
        spos = pay;
        svel = pay + 3 * n;
        smass = pay + 6 * n;
    }
    int64_t nocc = 0;
    rc = run_lengths(keys, n, pblocks, starts, &nocc);
    if (rc) return rc;
    // colour lists of occupied-block indices (p2g.rs:186-191), each block's
    // particles cut into items of at most kSplatChunk (dense blocks, e.g. a
    // droplet's, would otherwise serialise a whole colour pass on one work-group)
    std::vector<int32_t> hb((size_t)nocc), hs((size_t)nocc + 1);
    MSBG_HIP(hipMemcpy(hb.data(), pblocks, sizeof(int32_t) * nocc, hipMemcpyDeviceToHost));
    MSBG_HIP(hipMemcpy(hs.data(), starts, sizeof(int32_t) * (nocc + 1), hipMemcpyDeviceToHost));
    std::vector<SplatItem> order;
    int64_t cnt[8] = {0};
    for (int c = 0; c < 8; ++c)
        for (int64_t i = 0; i < nocc; ++i) {
            const int64_t b = hb[i];
            const int64_t bx = b % g->nx, by = (b / g->nx) % g->ny, bz = b / g->nxy;
            if ((int)((bx & 1) | ((by & 1) << 1) | ((bz & 1) << 2)) != c) continue;
            const int32_t p0 = hs[i], p1 = hs[i + 1];
            const int32_t shared = p1 - p0 > kSplatChunk ? 1 : 0;
            for (int32_t q = p0; q < p1; q += kSplatChunk) {
                order.push_back(SplatItem{(int32_t)i, q, std::min(p1, q + kSplatChunk), shared});
                ++cnt[c];
            }
        }
    SplatItem *d_order =
        (SplatItem *)workspace(3, sizeof(SplatItem) * std::max<size_t>(order.size(), 1), &rc);
    if (rc) return rc;
    MSBG_HIP(hipMemcpyAsync(d_order, order.data(), sizeof(SplatItem) * order.size(), hipMemcpyHostToDevice, stream()));
    const int m = (int)ceilf(r_p);
    const int msx = 16 + 2 * m;
    // both channels in one pass when two int64 windows and the residue list
    // fit (r_p <= 2), else a pass per channel with a float residue window
    const bool both = msx * msx * msx <= kSplatMaxCells;
    const size_t lds = both ? 2 * sizeof(unsigned long long) * msx * msx * msx + sizeof(uint2) * kResCap
                            : (sizeof(unsigned long long) + sizeof(float)) * msx * msx * msx;
    {
        int dev = 0, max_lds = 0;
        MSBG_HIP(hipGetDevice(&dev));
        MSBG_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        MSBG_CHECK((int64_t)lds <= max_lds, MSBG_EINVAL,
                   "particles_to_grid: the staging window for this r_p exceeds the workgroup LDS limit");
        if (lds > 65536) {
            const void *fns[5] = {(const void *)mac_splat_kernel<0, -1>, (const void *)mac_splat_kernel<1, -1>,
                                  (const void *)mac_splat_kernel<2, 0>, (const void *)mac_splat_kernel<2, 1>,
                                  (const void *)mac_splat_kernel<2, 2>};
            for (const void *f : fns)
                MSBG_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        }
    }
    const float inv_r2 = 1.0f / (r_p * r_p);
    for (int a = 0; a < 3; ++a) {
        int64_t off = 0;
        for (int c = 0; c < 8; ++c) {
            if (cnt[c] && both) {
                const dim3 grid((unsigned)cnt[c]), block(kSplatThreads);
                const GridView mv = mass[a]->view(), vv = vel[a]->view();
                if (a == 0)
                    mac_splat_kernel<2, 0><<<grid, block, lds, stream()>>>(mv, vv, a, d_order + off, pblocks, spos,
                                                                          svel, smass, r_p, inv_r2, m);
                else if (a == 1)
                    mac_splat_kernel<2, 1><<<grid, block, lds, stream()>>>(mv, vv, a, d_order + off, pblocks, spos,
                                                                          svel, smass, r_p, inv_r2, m);
                else
                    mac_splat_kernel<2, 2><<<grid, block, lds, stream()>>>(mv, vv, a, d_order + off, pblocks, spos,
                                                                          svel, smass, r_p, inv_r2, m);
            } else if (cnt[c]) {
                hipLaunchKernelGGL((mac_splat_kernel<0, -1>), dim3((unsigned)cnt[c]), dim3(kSplatThreads), lds,
                                   stream(), mass[a]->view(), vel[a]->view(), a, d_order + off, pblocks, spos, svel,
                                   smass, r_p, inv_r2, m);
                hipLaunchKernelGGL((mac_splat_kernel<1, -1>), dim3((unsigned)cnt[c]), dim3(kSplatThreads), lds,
                                   stream(), mass[a]->view(), vel[a]->view(), a, d_order + off, pblocks, spos, svel,
                                   smass, r_p, inv_r2, m);
            }
            off += cnt[c];
        }
        MSBG_HIP(hipGetLastError());
    }
    if (!act.empty())
        for (int a = 0; a < 3; ++a)
            hipLaunchKernelGGL(normalize_kernel, dim3((unsigned)act.size()), dim3(256), 0, stream(), vel[a]->view(),
                               mass[a]->view(), d_act);
    MSBG_HIP(hipGetLastError());
    MSBG_HIP(hipStreamSynchronize(stream()));
    return MSBG_OK;
}

int msbg_p2g_bspline(int order, msbg_grid *mass, msbg_grid *const *vel, const float *d_pos, const float *d_vel,
                     const float *d_mass, int64_t n, int64_t *n_active) {
    MSBG_CHECK(order == 1 || order == 2, MSBG_EINVAL, "p2g_bspline: order is 1 (trilinear) or 2 (quadratic)");
    MSBG_CHECK(mass && vel && n_active, MSBG_EINVAL, "p2g_bspline: null argument");
    msbg_grid *all[4] = {mass, vel[0], vel[1], vel[2]};
    int rc = check_same_dims(all, 4);
    if (rc) return rc;
    MSBG_CHECK(n >= 0 && n < ((int64_t)1 << 31) && (n == 0 || (d_pos && d_vel)), MSBG_EINVAL,
               "p2g_bspline: bad particle arrays");
    MSBG_CHECK(mass->sx <= 1024 && mass->sy <= 1024 && mass->sz <= 1024, MSBG_EINVAL,
               "p2g_bspline: the 30-bit Morton key needs a domain of at most 1024^3");
    // support cells floor(p)-1 .. floor(p)+1 lie in the footprint blocks of r_scan 2
    std::vector<int64_t> act;
    int32_t *d_act = nullptr;
    rc = active_list(mass, d_pos, n, 2.0f, act, &d_act);
    if (rc) return rc;
    rc = zero_blocks(all, 4, act);
    if (rc) return rc;
    *n_active = (int64_t)act.size();
    if (n == 0 || act.empty()) return MSBG_OK;
    // Morton sort (stable) and the sorted payload
    uint32_t *keys = (uint32_t *)workspace(4, sizeof(uint32_t) * 2 * n, &rc);
    if (rc) return rc;
    uint32_t *perm = keys + n;
    float4 *rec = (float4 *)workspace(22, sizeof(float4) * 2 * n, &rc);
    if (rc) return rc;
    rc = morton_sort_pack_async(d_pos, d_vel, d_mass, n, keys, perm, rec);
    if (rc) return rc;
    float4 *pay = (float4 *)workspace(7, sizeof(float4) * 2 * n, &rc);
    if (rc) return rc;
    float4 *pm = pay, *pu = pay + n;
    hipLaunchKernelGGL(gather_records_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream(), rec, perm, n, pm, pu);
    MSBG_HIP(hipGetLastError());
    // per-voxel ranges of the sorted particles, indexed like the mass slab
    const int64_t ncell = mass->n_slots * 4096;
    uint32_t *vr = (uint32_t *)workspace(8, sizeof(uint32_t) * 2 * ncell, &rc);
    if (rc) return rc;
    uint32_t *vbeg = vr, *vend = vr + ncell;
    MSBG_HIP(hipMemsetAsync(vr, 0, sizeof(uint32_t) * 2 * ncell, stream()));
    hipLaunchKernelGGL(voxel_ranges_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream(), keys, pm, n, mass->view(),
                       vbeg, vend);
    MSBG_HIP(hipGetLastError());
    const dim3 grid((unsigned)act.size()), block(256);
    if (order == 1)
        hipLaunchKernelGGL(p2g_bspline_kernel<1>, grid, block, 0, stream(), mass->view(), vel[0]->view(), vel[1]->view(),
                           vel[2]->view(), d_act, vbeg, vend, pm, pu);
    else
        hipLaunchKernelGGL(p2g_bspline_kernel<2>, grid, block, 0, stream(), mass->view(), vel[0]->view(), vel[1]->view(),
                           vel[2]->view(), d_act, vbeg, vend, pm, pu);
    MSBG_HIP(hipGetLastError());
    MSBG_HIP(hipStreamSynchronize(stream()));
    return MSBG_OK;
}

int msbg_g2p_bspline(int order, msbg_grid *const *vel, const float *d_pos, int64_t n, float *d_out) {
    MSBG_CHECK(order == 1 || order == 2, MSBG_EINVAL, "g2p_bspline: order is 1 (trilinear) or 2 (quadratic)");
    MSBG_CHECK(vel, MSBG_EINVAL, "g2p_bspline: null grids");
    int rc = check_same_dims(vel, 3);
    if (rc) return rc;
    MSBG_CHECK(n >= 0 && (n == 0 || (d_pos && d_out)), MSBG_EINVAL, "g2p_bspline: null argument");
    MSBG_CHECK(n < ((int64_t)1 << 31), MSBG_EINVAL, "g2p_bspline: n must be below 2^31");
    if (n == 0) return MSBG_OK;
    uint32_t *keys = (uint32_t *)workspace(9, sizeof(uint32_t) * 2 * n, &rc);
    if (rc) return rc;
    uint32_t *visit = keys + n;
    rc = morton_sort_async(d_pos, n, keys, visit);
    if (rc) return rc;
    // one lane per particle (n < 2^31, so the chunk count fits the grid)
    const dim3 chunks((unsigned)((n + kG2pThreads - 1) / kG2pThreads)), lanes(kG2pThreads);
    if (order == 1)
        hipLaunchKernelGGL(g2p_bspline_kernel<1>, chunks, lanes, 0, stream(), vel[0]->view(), vel[1]->view(),
                           vel[2]->view(), d_pos, n, visit, d_out);
    else
        hipLaunchKernelGGL(g2p_bspline_kernel<2>, chunks, lanes, 0, stream(), vel[0]->view(), vel[1]->view(),
                           vel[2]->view(), d_pos, n, visit, d_out);
    MSBG_HIP(hipGetLastError());
    MSBG_HIP(hipStreamSynchronize(stream()));
    return MSBG_OK;
}

}  // extern "C"

// ===========================================================================
// MacGrid::divergence -- fluid/mac.rs:137-171 (each cell at its layout index;
// the reference's running index coincides for whole blocks).
// ===========================================================================
namespace msbg {
namespace {

__device__ __forceinline__ float voxel_at(const GridView &g, int x, int y, int z) {
    const int64_t bid = (x >> 4) + (int64_t)(y >> 4) * g.nx + (int64_t)(z >> 4) * g.nx * g.ny;
    return g.slab[((int64_t)g.b2s[bid] << 12) + ((x & 15) | ((y & 15) << 4) | ((z & 15) << 8))];
}

// The divergence and projection kernels run 16 work-groups of 256 cells per
// listed block, so the block's slot lookups (its own and its neighbours')
// are uniform over a work-group and go out once, together; each lane then
// issues its value loads together.  The values are those voxel_at reads.
__global__ __launch_bounds__(256) void divergence_kernel(GridView u, GridView v, GridView w, GridView d,
                                                         const int32_t *__restrict__ bids, int64_t n) {
    const int64_t k = (int64_t)(blockIdx.x >> 4);
    if (k >= n) return;
    const int vid = (int)((blockIdx.x & 15) << 8) | (int)threadIdx.x;
    const int64_t bid = bids[k];
    const int slot = d.b2s[bid];
    if (slot < 2) return;
    const int64_t nxy = (int64_t)d.nx * d.ny;
    const int bx = (int)(bid % d.nx), by = (int)((bid / d.nx) % d.ny), bz = (int)(bid / nxy);
    const int lx = vid & 15, ly = (vid >> 4) & 15, lz = vid >> 8;
    const int x = bx * 16 + lx, y = by * 16 + ly, z = bz * 16 + lz;
    if (x >= d.sx || y >= d.sy || z >= d.sz) return;
    // this block's slots in the face grids and its -x / -y / -z neighbours' (itself where none exists)
    const int64_t su = (int64_t)u.b2s[bid] << 12, sv = (int64_t)v.b2s[bid] << 12, sw = (int64_t)w.b2s[bid] << 12;
    const int64_t sux = (int64_t)u.b2s[bx > 0 ? bid - 1 : bid] << 12;
    const int64_t svy = (int64_t)v.b2s[by > 0 ? bid - d.nx : bid] << 12;
    const int64_t swz = (int64_t)w.b2s[bz > 0 ? bid - nxy : bid] << 12;
    const float ux1 = u.slab[su + vid], vy1 = v.slab[sv + vid], wz1 = w.slab[sw + vid];
    const float um = u.slab[lx > 0 ? su + vid - 1 : sux + vid + 15];
    const float vm = v.slab[ly > 0 ? sv + vid - 16 : svy + vid + 240];
    const float wm = w.slab[lz > 0 ? sw + vid - 256 : swz + vid + 3840];
    const float ux0 = x == 0 ? 0.0f : um, vy0 = y == 0 ? 0.0f : vm, wz0 = z == 0 ? 0.0f : wm;
    d.slab[((int64_t)slot << 12) + vid] = ((ux1 - ux0) + (vy1 - vy0)) + (wz1 - wz0);
}

// Pressure projection of the face velocities: face a of cell c takes
// u_a(c) - scale (p(c + e_a) - p(c)), the neighbour outside the domain being
// the cell itself; one thread per cell, the three faces it owns.  LO: p is
// the compensated pair p + plo (plo over p's slot layout), and the gradient
// is (p(c + e_a) - p(c)) + (plo(c + e_a) - plo(c)).
template <bool LO>
__global__ __launch_bounds__(256) void subtract_gradient_kernel(GridView u, GridView v, GridView w, GridView p,
                                                                const float *__restrict__ plo,
                                                                const int32_t *__restrict__ bids, int64_t n,
                                                                float scale) {
    const int64_t k = (int64_t)(blockIdx.x >> 4);
    if (k >= n) return;
    const int vid = (int)((blockIdx.x & 15) << 8) | (int)threadIdx.x;
    const int64_t bid = bids[k];
    const int64_t nxy = (int64_t)p.nx * p.ny;
    const int bx = (int)(bid % p.nx), by = (int)((bid / p.nx) % p.ny), bz = (int)(bid / nxy);
    const int lx = vid & 15, ly = (vid >> 4) & 15, lz = vid >> 8;
    const int x = bx * 16 + lx, y = by * 16 + ly, z = bz * 16 + lz;
    if (x >= p.sx || y >= p.sy || z >= p.sz) return;
    // p's slots of this block and of its +x / +y / +z neighbours (itself where none exists)
    const int64_t sp = (int64_t)p.b2s[bid] << 12;
    const int64_t spx = (int64_t)p.b2s[bx + 1 < p.nx ? bid + 1 : bid] << 12;
    const int64_t spy = (int64_t)p.b2s[by + 1 < p.ny ? bid + p.nx : bid] << 12;
    const int64_t spz = (int64_t)p.b2s[bz + 1 < p.nz ? bid + nxy : bid] << 12;
    const int64_t cx = lx < 15 ? sp + vid + 1 : spx + vid - 15, cy = ly < 15 ? sp + vid + 16 : spy + vid - 240,
                  cz = lz < 15 ? sp + vid + 256 : spz + vid - 3840;
    const bool hx = x + 1 < p.sx, hy = y + 1 < p.sy, hz = z + 1 < p.sz;  // the neighbour is in the domain
    const float pc = p.slab[sp + vid];
    const float px = p.slab[hx ? cx : sp + vid], py = p.slab[hy ? cy : sp + vid], pz = p.slab[hz ? cz : sp + vid];
    const float pn[3] = {hx ? px : pc, hy ? py : pc, hz ? pz : pc};
    float lc = 0.0f, ln[3] = {0.0f, 0.0f, 0.0f};
    if constexpr (LO) {  // plo has p's slot layout
        lc = plo[sp + vid];
        const float qx = plo[hx ? cx : sp + vid], qy = plo[hy ? cy : sp + vid], qz = plo[hz ? cz : sp + vid];
        ln[0] = hx ? qx : lc;
        ln[1] = hy ? qy : lc;
        ln[2] = hz ? qz : lc;
    }
    const GridView ch[3] = {u, v, w};
    int slots[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) slots[a] = ch[a].b2s[bid];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (slots[a] < 2) continue;
        float *f = ch[a].slab + ((int64_t)slots[a] << 12) + vid;
        float g = pn[a] - pc;
        if constexpr (LO) g = g + (ln[a] - lc);
        const float t = scale * g;
        *f = *f - t;
    }
}

}  // namespace
}  // namespace msbg

namespace msbg {
namespace {

int subtract_gradient(msbg_grid *const *vel, const msbg_grid *p_layout, const float *d_p, const float *d_plo,
                      const int64_t *active, int64_t n, float scale) {
    MSBG_CHECK(vel && vel[0] && vel[1] && vel[2] && p_layout && d_p, MSBG_EINVAL, "subtract gradient: null argument");
    MSBG_CHECK(n >= 0 && (n == 0 || active), MSBG_EINVAL, "subtract gradient: bad active list");
    for (int a = 0; a < 3; ++a)
        MSBG_CHECK(vel[a]->L == 4 && vel[a]->sx == p_layout->sx && vel[a]->sy == p_layout->sy &&
                       vel[a]->sz == p_layout->sz,
                   MSBG_EINVAL, "subtract gradient: grids must share dimensions and 16^3 blocks");
    MSBG_CHECK(p_layout->L == 4, MSBG_EINVAL, "subtract gradient: 16^3-block pressure layout required");
    if (!n) return MSBG_OK;
    for (int64_t k = 0; k < n; ++k)
        MSBG_CHECK(active[k] >= 0 && active[k] < p_layout->n_blocks, MSBG_EINVAL, "subtract gradient: bid out of range");
    int rc;
    int32_t *d_bids = (int32_t *)workspace(19, sizeof(int32_t) * n, &rc);
    if (rc) return rc;
    std::vector<int32_t> h(active, active + n);
    MSBG_HIP(hipMemcpyAsync(d_bids, h.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, stream()));
    GridView pv = p_layout->view();
    pv.slab = const_cast<float *>(d_p);
    const int grid = (int)((n * 4096 + 255) / 256);
    if (d_plo)
        subtract_gradient_kernel<true><<<grid, 256, 0, stream()>>>(vel[0]->view(), vel[1]->view(), vel[2]->view(), pv,
                                                                    d_plo, d_bids, n, scale);
    else
        subtract_gradient_kernel<false><<<grid, 256, 0, stream()>>>(vel[0]->view(), vel[1]->view(), vel[2]->view(),
                                                                     pv, nullptr, d_bids, n, scale);
    MSBG_HIP(hipGetLastError());
    MSBG_HIP(hipStreamSynchronize(stream()));
    return MSBG_OK;
}

}  // namespace
}  // namespace msbg

extern "C" int msbg_mac_subtract_gradient(msbg_grid *const *vel, const msbg_grid *p_layout, const float *d_p,
                                          const int64_t *active, int64_t n, float scale) {
    return msbg::subtract_gradient(vel, p_layout, d_p, nullptr, active, n, scale);
}

extern "C" int msbg_mac_subtract_gradient_pair(msbg_grid *const *vel, const msbg_grid *p_layout, const float *d_p,
                                               const float *d_p_lo, const int64_t *active, int64_t n, float scale) {
    MSBG_CHECK(d_p_lo, MSBG_EINVAL, "subtract gradient: null low part");
    return msbg::subtract_gradient(vel, p_layout, d_p, d_p_lo, active, n, scale);
}

extern "C" int msbg_mac_divergence(msbg_grid *const *vel, msbg_grid *dst, const int64_t *active, int64_t n) {
    MSBG_CHECK(vel && vel[0] && vel[1] && vel[2] && dst, MSBG_EINVAL, "divergence: null grid");
    MSBG_CHECK(n >= 0 && (n == 0 || active), MSBG_EINVAL, "divergence: bad active list");
    for (int a = 0; a < 3; ++a)
        MSBG_CHECK(vel[a]->L == 4 && vel[a]->sx == dst->sx && vel[a]->sy == dst->sy && vel[a]->sz == dst->sz,
                   MSBG_EINVAL, "divergence: grids must share dimensions and 16^3 blocks");
    MSBG_CHECK(dst->L == 4, MSBG_EINVAL, "divergence: 16^3-block destination required");
    if (!n) return MSBG_OK;
    int rc = msbg_grid_ensure_blocks(dst, active, n);  // dst.ensure_blocks_parallel(active), mac.rs:142
    if (rc) return rc;
    int32_t *d_bids = (int32_t *)workspace(15, sizeof(int32_t) * n, &rc);
    if (rc) return rc;
    std::vector<int32_t> h(active, active + n);
    MSBG_HIP(hipMemcpyAsync(d_bids, h.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, stream()));
    divergence_kernel<<<(int)((n * 4096 + 255) / 256), 256, 0, stream()>>>(vel[0]->view(), vel[1]->view(),
                                                                             vel[2]->view(), dst->view(), d_bids, n);
    MSBG_HIP(hipGetLastError());
    MSBG_HIP(hipStreamSynchronize(stream()));
    return MSBG_OK;
}

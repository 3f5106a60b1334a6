// Surface-reconstruction tail: instance placement of a base point cloud
// (particles/sort.rs `place`) and the distance-to-density finalize of a
// Density (u16) slab (particles/finalize.rs).  The splat in between is
// msbg_density_splat, the smoothing after it msbg_sweep_density.
#include <algorithm>
#include <cmath>

#include "internal.hpp"
#include "primitives.hpp"

namespace msbg {
namespace {

constexpr int kPlaceThreads = 256;

struct PlaceArgs {
    const float *base;
    int64_t n_base;
    float bmin[3];
    float smax, base_scale, inst_scale;
    float lo[3], hi[3];
    int32_t s[3];
    int32_t nx, nxy;
};

// Origin of instance i: 0.2 smax + ((0.6 smax) base_scale) (b_i - bbox), each
// operation rounded to f32 in that order (sort.rs:71-75).
__device__ __forceinline__ void instance_origin(const PlaceArgs &a, int64_t i, float *o) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float t0 = 0.2f * a.smax;
        float t1 = 0.6f * a.smax;
        t1 = t1 * a.base_scale;
        t1 = t1 * (a.base[3 * i + k] - a.bmin[k]);
        o[k] = t0 + t1;
    }
}

// Base point j of an instance: kept when inside the domain bounds and its
// truncation lies in the grid (sort.rs:79-101).
__device__ __forceinline__ bool place_point(const PlaceArgs &a, const float *o, int64_t j, float *p, int32_t &bid) {
    bool in = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        p[k] = o[k] + a.inst_scale * (a.base[3 * j + k] - a.bmin[k]);
        in = in && p[k] >= a.lo[k] && p[k] <= a.hi[k];
    }
    if (!in) return false;
    int32_t ip[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        ip[k] = (int32_t)p[k];
        if (ip[k] < 0 || ip[k] >= a.s[k]) return false;
    }
    bid = (ip[0] >> 4) + (ip[1] >> 4) * a.nx + (ip[2] >> 4) * a.nxy;
    return true;
}

// Pass 1: kept particles of instance blockIdx.x.
__global__ __launch_bounds__(kPlaceThreads) void place_count_kernel(PlaceArgs a, uint32_t *counts) {
    const int64_t i = blockIdx.x;
    float o[3];
    instance_origin(a, i, o);
    uint32_t c = 0;
    for (int64_t j = threadIdx.x; j < a.n_base; j += kPlaceThreads) {
        float p[3];
        int32_t bid;
        c += place_point(a, o, j, p, bid) ? 1u : 0u;
    }
    for (int d = 32; d > 0; d >>= 1) c += __shfl_xor(c, d);
    __shared__ uint32_t s_w[kPlaceThreads / 64];
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) counts[i] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// Pass 2: instance blockIdx.x writes its kept particles from offsets[i] on,
// in base-point order (wave ballots rank them inside a 256-point chunk).
__global__ __launch_bounds__(kPlaceThreads) void place_write_kernel(PlaceArgs a, const uint32_t *offsets,
                                                                    float *pos, int32_t *bids) {
    const int64_t i = blockIdx.x;
    float o[3];
    instance_origin(a, i, o);
    __shared__ uint32_t s_w[kPlaceThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int64_t run = offsets[i];
    for (int64_t j0 = 0; j0 < a.n_base; j0 += kPlaceThreads) {
        const int64_t j = j0 + threadIdx.x;
        float p[3];
        int32_t bid = 0;
        const bool keep = j < a.n_base && place_point(a, o, j, p, bid);
        const uint64_t m = __ballot(keep);
        const uint32_t below = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) s_w[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (int w = 0; w < kPlaceThreads / 64; ++w) {
            before += w < wave ? s_w[w] : 0u;
            total += s_w[w];
        }
        if (keep) {
            const int64_t k = run + before + below;
            pos[3 * k] = p[0];
            pos[3 * k + 1] = p[1];
            pos[3 * k + 2] = p[2];
            bids[k] = bid;
        }
        run += total;
        __syncthreads();
    }
}

// finalize -- finalize.rs:47-69 on 8 cells per thread (one 16-byte row
// vector); 512 threads cover a block.
__global__ __launch_bounds__(256) void finalize_kernel(uint16_t *slab, const int32_t *b2s, const int32_t *active,
                                                       float r_particle, float dist_sq_max, float denom) {
    const int64_t blk = blockIdx.x >> 1;
    const int32_t slot = b2s[active[blk]];
    if (slot < 2) return;  // no value block: value_block_ptr_mut is null
    uint16_t *d = slab + (int64_t)slot * 4096 + (int64_t)((blockIdx.x & 1) * 256 + threadIdx.x) * 8;
    const uint4 q = *reinterpret_cast<const uint4 *>(d);
    const uint32_t w[4] = {q.x, q.y, q.z, q.w};
    uint32_t o[4];
    const float inv = 1.0f / 65535.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t u = (w[k >> 1] >> (16 * (k & 1))) & 0xffffu;
        const float ratio = (float)u * inv;
        const bool touched = ratio < 0.999f;
        const float dist_sq = ratio * dist_sq_max;
        const float dist = __builtin_sqrtf(dist_sq) - r_particle;
        float t = (dist + r_particle) / denom;
        t = fminf(fmaxf(t, 0.0f), 1.0f);
        const float f = 1.0f - t;
        const float out = touched ? f : 0.0f;
        const uint32_t c = (uint32_t)(int32_t)__builtin_roundf(out * 65535.0f);
        if (k & 1)
            o[k >> 1] |= c << 16;
        else
            o[k >> 1] = c;
    }
    *reinterpret_cast<uint4 *>(d) = make_uint4(o[0], o[1], o[2], o[3]);
}

// DomainBounds::new -- particles/mod.rs:180-194 (24-bit fixed point).
float align24(double x) {
    const double scale = 16777215.0 / 32768.0;
    return (float)(std::floor(scale * x + 0.5) / scale);
}

}  // namespace
}  // namespace msbg

using namespace msbg;

extern "C" {

int msbg_surface_place(const float *d_base, int64_t n_base, const float *bbox_min, float span_max, int64_t sx,
                       int64_t sy, int64_t sz, int64_t n_instances, float instance_scale_factor, float *d_pos_out,
                       int32_t *d_bids_out, int64_t capacity, int64_t *n_out) {
    MSBG_CHECK(bbox_min && n_out, MSBG_EINVAL, "surface place: null argument");
    MSBG_CHECK(n_base >= 0 && (n_base == 0 || d_base), MSBG_EINVAL, "surface place: bad base cloud");
    MSBG_CHECK(sx > 0 && sy > 0 && sz > 0 && sx < (1 << 24) && sy < (1 << 24) && sz < (1 << 24), MSBG_EINVAL,
               "surface place: bad domain");
    MSBG_CHECK(span_max > 0.0f && std::isfinite(span_max), MSBG_EINVAL,
               "surface place: degenerate point cloud (zero span)");
    MSBG_CHECK(n_instances >= 0 && n_instances <= n_base, MSBG_EINVAL,
               "surface place: n_instances must be 0 (one per base point) or at most the base point count");
    MSBG_CHECK(capacity >= 0 && (capacity == 0 || (d_pos_out && d_bids_out)), MSBG_EINVAL,
               "surface place: bad output buffers");
    const int64_t n_inst = n_instances == 0 ? n_base : n_instances;
    MSBG_CHECK(n_inst < ((int64_t)1 << 31), MSBG_EINVAL, "surface place: too many instances");
    MSBG_CHECK(n_inst * n_base < ((int64_t)1 << 32), MSBG_EINVAL,
               "surface place: n_instances x n_base must stay below 2^32 candidate particles");
    *n_out = 0;
    if (n_inst == 0) return MSBG_OK;
    PlaceArgs a;
    a.base = d_base;
    a.n_base = n_base;
    for (int k = 0; k < 3; ++k) a.bmin[k] = bbox_min[k];
    const int64_t s[3] = {sx, sy, sz};
    a.smax = (float)std::max(sx, std::max(sy, sz));
    a.base_scale = 1.0f / span_max;
    a.inst_scale = ((float)std::min(sx, std::min(sy, sz)) * instance_scale_factor) * a.base_scale;
    for (int k = 0; k < 3; ++k) {
        a.lo[k] = align24(0.005);
        a.hi[k] = align24((double)s[k] - 0.005);
        a.s[k] = (int32_t)s[k];
    }
    a.nx = (int32_t)((sx + 15) >> 4);
    a.nxy = a.nx * (int32_t)((sy + 15) >> 4);
    int rc;
    uint32_t *counts = (uint32_t *)workspace(16, sizeof(uint32_t) * (size_t)n_inst, &rc);
    if (rc) return rc;
    hipLaunchKernelGGL(place_count_kernel, dim3((unsigned)n_inst), dim3(kPlaceThreads), 0, stream(), a, counts);
    MSBG_HIP(hipGetLastError());
    uint32_t total = 0;
    rc = scan_exclusive_u32(counts, counts, n_inst, &total);
    if (rc) return rc;
    *n_out = total;
    if (capacity == 0 && !d_pos_out && !d_bids_out) return MSBG_OK;  // count only
    MSBG_CHECK((int64_t)total <= capacity, MSBG_ENOMEM,
               "surface place: more placed particles than the output capacity (see *n_out)");
    hipLaunchKernelGGL(place_write_kernel, dim3((unsigned)n_inst), dim3(kPlaceThreads), 0, stream(), a, counts,
                       d_pos_out, d_bids_out);
    MSBG_HIP(hipGetLastError());
    return MSBG_OK;
}

int msbg_density_finalize(const msbg_grid *layout, uint16_t *d_slab16, int64_t slab_slots, const int64_t *active,
                          int64_t n, float r_particle, float nb_dist) {
    MSBG_CHECK(layout && d_slab16 && (n == 0 || active), MSBG_EINVAL, "density finalize: null argument");
    MSBG_CHECK(layout->L == 4, MSBG_EINVAL, "density finalize: 16^3-block grid required");
    MSBG_CHECK(layout->n_slots <= slab_slots, MSBG_EINVAL, "density finalize: the u16 slab has fewer slots than the grid");
    MSBG_CHECK(n >= 0 && n < ((int64_t)1 << 30), MSBG_EINVAL, "density finalize: bad active count");
    if (n == 0) return MSBG_OK;
    std::vector<int32_t> act((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        MSBG_CHECK(active[i] >= 0 && active[i] < layout->n_blocks, MSBG_EINVAL, "density finalize: bid out of range");
        act[(size_t)i] = (int32_t)active[i];
    }
    int rc;
    int32_t *d_act = (int32_t *)workspace(17, sizeof(int32_t) * (size_t)n, &rc);
    if (rc) return rc;
    MSBG_HIP(hipMemcpyAsync(d_act, act.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, stream()));
    const float r_scan = r_particle + nb_dist;  // SurfaceConfig::r_scan, particles/mod.rs:160-162
    const float dist_sq_max = r_scan * r_scan;
    const float denom = 2.0f * r_particle + nb_dist;
    hipLaunchKernelGGL(finalize_kernel, dim3((unsigned)(2 * n)), dim3(256), 0, stream(), d_slab16, layout->d_b2s,
                       d_act, r_particle, dist_sq_max, denom);
    MSBG_HIP(hipGetLastError());
    MSBG_HIP(hipStreamSynchronize(stream()));  // `act` is pageable host memory
    return MSBG_OK;
}

}  // extern "C"

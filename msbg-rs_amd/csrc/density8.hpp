// Density8 (u8) storage of the smoother (channel.rs:55-59,155-189): loaded
// linearly as code * (1/255) (the Density8 Dequant, math/sample.rs:64-69) and
// stored sqrt-compressed with stochastic rounding (StoreBack<Density8>,
// math/stencil.rs:176-192): vf = sqrt(v) * 255, code = floor(vf) + (u < vf -
// floor(vf)) for a uniform draw u in [0, 1).
//
// The reference draws u from a per-thread-slot lane-parallel generator
// (SimdRng, math/stencil.rs:101-132) whose stream depends on the SIMD width
// and on the work-stealing schedule.  Here u is a function of (seed, sweep,
// bid, vid): murmur3's 32-bit finaliser (the mixer SimdRng::seed uses) chained
// over the sweep index, the block id and the voxel, one Lehmer step
// (x 16807, SimdRng::next) and the top 23 bits as a float in [0, 1) -- so every
// stored code is reproducible on any device and schedule.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace msbg {

__host__ __device__ __forceinline__ uint32_t fmix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

// The per-sweep part of the stream (one value per work-group).
__host__ __device__ __forceinline__ uint32_t sr_sweep_key(uint32_t seed, uint32_t sweep) {
    return fmix32(seed ^ (sweep * 0x9e3779b9u));
}

// Uniform [0, 1) draw of voxel vid of block bid in the sweep keyed `key`.
__host__ __device__ __forceinline__ float sr_uniform(uint32_t key, uint32_t bid, uint32_t vid) {
    uint32_t h = fmix32(fmix32(key ^ bid) ^ (vid * 0x9e3779b9u));
    if (h == 0u) h = 0xdeadbeefu;
    const uint32_t bits = ((h * 16807u) >> 9) | 0x3f800000u;
    float f;
    __builtin_memcpy(&f, &bits, 4);
    return f - 1.0f;
}

// The stochastic sqrt-compressed code of v (v in [0, 1]: the result is
// floor(vf) or floor(vf) + 1, vf = sqrt(v) * 255); outside [0, 1] it
// saturates to [0, 255] (NaN -> 0).
__host__ __device__ __forceinline__ uint8_t sr_quant8(float v, float u) {
    const float vf = __builtin_sqrtf(v) * 255.0f;
    const float fl = __builtin_floorf(vf);
    const float r = u < vf - fl ? fl + 1.0f : fl;
    return r >= 255.0f ? (uint8_t)255 : (r > 0.0f ? (uint8_t)(int)r : (uint8_t)0);
}

}  // namespace msbg

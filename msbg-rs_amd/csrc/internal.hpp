// Internal definitions shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/msbg_hip.h"

namespace msbg {

// Records `msg` as the thread's last error and returns `code`.
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);
hipStream_t stream();

#define MSBG_HIP(call)                                          \
    do {                                                        \
        hipError_t e_ = (call);                                 \
        if (e_ != hipSuccess) return ::msbg::hip_fail(e_, #call); \
    } while (0)

#define MSBG_CHECK(cond, code, msg)                      \
    do {                                                 \
        if (!(cond)) return ::msbg::fail((code), (msg)); \
    } while (0)

// A host-visible description of one grid level, passed to kernels by value.
struct GridView {
    float *slab;       // device, n_slots * N
    const int32_t *b2s;  // device, n_blocks
    int32_t sx, sy, sz;  // voxels
    int32_t nx, ny, nz;  // blocks
    int32_t L;           // log2 block edge
};

}  // namespace msbg

struct msbg_grid {
    int64_t sx, sy, sz, nx, ny, nz, nxy, n_blocks;
    int L, B;
    int64_t N;  // B^3
    int64_t capacity_slots;  // including the 2 dummies
    int64_t n_slots;
    float empty_value, full_value;
    std::vector<int32_t> b2s_host;
    float *d_slab = nullptr;
    int32_t *d_b2s = nullptr;

    msbg::GridView view() const {
        msbg::GridView v;
        v.slab = d_slab;
        v.b2s = d_b2s;
        v.sx = (int32_t)sx;
        v.sy = (int32_t)sy;
        v.sz = (int32_t)sz;
        v.nx = (int32_t)nx;
        v.ny = (int32_t)ny;
        v.nz = (int32_t)nz;
        v.L = L;
        return v;
    }
};

// ---------------------------------------------------------------------------
// Device helpers: boundary resolution (math/boundary.rs:35-56).
// ---------------------------------------------------------------------------
namespace msbg {

// Work-group index -> work item, XCD-aware: the dispatcher hands work-group i
// to XCD i % 8, so give each XCD a contiguous run of the work list (blocks in
// Morton order), letting neighbouring blocks share that XCD's L2.  A bijection
// on [0, n) whatever the dispatch order is.
__device__ __forceinline__ int xcd_remap(int i, int n) {
    constexpr int kXcds = 8;
    const int q = n / kXcds, r = n % kXcds;
    const int xcd = i % kXcds, k = i / kXcds;
    return xcd * q + (xcd < r ? xcd : r) + k;
}

__device__ __forceinline__ int reflect_axis(int g, int max) {
    int period = 2 * (max + 1);
    int r = g % period;
    if (r < 0) r += period;
    return r > max ? period - 1 - r : r;
}

// Returns false only for Dirichlet out of range.
__device__ __forceinline__ bool resolve_axis(int g, int max, int bc, int &out) {
    if (g >= 0 && g <= max) {
        out = g;
        return true;
    }
    if (bc == MSBG_BC_CLAMP) {
        out = g < 0 ? 0 : max;
        return true;
    }
    if (bc == MSBG_BC_NEUMANN) {
        out = reflect_axis(g, max);
        return true;
    }
    return false;
}

}  // namespace msbg

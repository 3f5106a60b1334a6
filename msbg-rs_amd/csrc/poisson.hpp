// Internal definitions of the Poisson band (msbg_poisson) shared by the
// relaxation / solve unit and the learned-correction unit.
#pragma once

#include "internal.hpp"

namespace msbg {

using i64 = int64_t;

namespace mg {
struct Hierarchy;
}

// One listed value block: its slot, voxel origin and the slots of its six
// face neighbours (-x, +x, -y, +y, -z, +z; unused where the face is on the
// domain boundary).
struct PBlock {
    int32_t slot, ox, oy, oz;
    int32_t nb[6];
};

struct Dom {
    int32_t sx, sy, sz;
};

// Element access of the three storage types: ld widens to FP32, cv narrows
// with round-to-nearest-even.
template <int ST>
struct Sto;
template <>
struct Sto<MSBG_STORAGE_F32> {
    using T = float;
    static __device__ __forceinline__ float ld(const T *p, i64 i) { return p[i]; }
    static __device__ __forceinline__ void ld4(const T *p, i64 i, float *v) {  // i: multiple of 4
        const float4 q = *reinterpret_cast<const float4 *>(p + i);
        v[0] = q.x;
        v[1] = q.y;
        v[2] = q.z;
        v[3] = q.w;
    }
    static __device__ __forceinline__ T cv(float v) { return v; }
};
template <>
struct Sto<MSBG_STORAGE_F16> {
    using T = uint16_t;
    static __device__ __forceinline__ float ld(const T *p, i64 i) { return (float)__builtin_bit_cast(_Float16, p[i]); }
    static __device__ __forceinline__ float lo(uint32_t w) {
        return (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xffffu));
    }
    static __device__ __forceinline__ float hi(uint32_t w) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16)); }
    static __device__ __forceinline__ void ld4(const T *p, i64 i, float *v) {
        const uint2 q = *reinterpret_cast<const uint2 *>(p + i);
        v[0] = lo(q.x);
        v[1] = hi(q.x);
        v[2] = lo(q.y);
        v[3] = hi(q.y);
    }
    // The FP32 value is materialised before the RNE narrowing: without the
    // register barrier the compiler fuses fmaf + cvt into one mixed-precision
    // FMA (v_fma_mixlo_f16), which rounds once to f16 instead of f32 then f16.
    static __device__ __forceinline__ T cv(float v) {
        asm volatile("" : "+v"(v));
        return __builtin_bit_cast(uint16_t, (_Float16)v);
    }
};
template <>
struct Sto<MSBG_STORAGE_BF16> {
    using T = uint16_t;
    static __device__ __forceinline__ float ld(const T *p, i64 i) { return __uint_as_float((uint32_t)p[i] << 16); }
    static __device__ __forceinline__ float lo(uint32_t w) { return __uint_as_float(w << 16); }
    static __device__ __forceinline__ float hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }
    static __device__ __forceinline__ void ld4(const T *p, i64 i, float *v) {
        const uint2 q = *reinterpret_cast<const uint2 *>(p + i);
        v[0] = lo(q.x);
        v[1] = hi(q.x);
        v[2] = lo(q.y);
        v[3] = hi(q.y);
    }
    static __device__ __forceinline__ T cv(float v) {
        uint32_t u = __float_as_uint(v);
        if ((u & 0x7fffffffu) > 0x7f800000u) return (T)((u >> 16) | 0x40u);
        u += 0x7fffu + ((u >> 16) & 1u);
        return (T)(u >> 16);
    }
};

// Per-block kernels of 256 threads; the cell-wise ones (correction, dots) take
// kSlices work-groups per block, 4 z-slices each, 16 x 16 threads.
constexpr int kThreads = 256;
constexpr int kSlices = 4;

}  // namespace msbg

struct msbg_poisson {
    msbg::Dom dom{};
    int64_t n_slots = 0, n_active = 0, n_cells = 0;
    int64_t nx = 0, ny = 0, nz = 0, n_blocks = 0;
    int n_wg = 0;                     // work-groups of the kSlices-per-block kernels
    int zcol[3] = {1, 1, 1};          // column depth of the relaxation kernels, per storage
    msbg::PBlock *d_blocks = nullptr;
    std::vector<int64_t> bids;       // the listed block ids, in list order
    msbg::mg::Hierarchy *mg = nullptr;  // multigrid levels, built by the first multigrid solve
    int32_t *d_b2s = nullptr;        // copy of the layout grid's bid -> slot table
    uint8_t *d_listed = nullptr;     // per block: 1 if listed
    double2 *d_partial = nullptr;
    double *d_norm = nullptr;   // [2]
    float *d_scale = nullptr;   // [2]: 2^k ~ rms(r), 2^-k
    double *h_norm = nullptr;   // pinned [2]
    float *d_r = nullptr;       // FP32 residual, n_slots * 4096
    void *d_e = nullptr, *d_rt = nullptr;  // inner fields (2 or 4 B elements)
    int inner_bytes = 0;
    bool e_zeroed = false;
    // learned correction scratch
    float *d_delta = nullptr, *d_xtry = nullptr;
    double *d_dots = nullptr;  // [2]: r . A delta, |A delta|^2
    // graph of one check period, keyed by its arguments; g_done / j_done are
    // recorded after each launch, and an exec is destroyed only once its
    // event has completed
    hipStream_t cap_stream = nullptr;
    hipEvent_t g_done = nullptr, j_done = nullptr;
    hipGraphExec_t exec = nullptr;
    int g_mode = -1, g_storage = -1, g_sweeps = -1;
    float g_omega = 0.0f;
    const void *g_x = nullptr, *g_lo = nullptr, *g_b = nullptr;
    // graph of a run of Jacobi sweeps, keyed by its arguments
    hipGraphExec_t jexec = nullptr;
    int j_storage = -1, j_sweeps = -1;
    float j_omega = 0.0f;
    const void *j_x = nullptr, *j_t = nullptr, *j_b = nullptr;
};

namespace msbg {
namespace pband {

// r = b - A x over the listed cells (r optional) and the FP64 norms into
// p->d_norm, the refinement scale into p->d_scale.
void residual(msbg_poisson *p, hipStream_t s, int storage, const void *x, const void *b, float *r);
// Copies p->d_norm to the host (synchronizes the library stream).
int read_norms(msbg_poisson *p, double *r2, double *b2);
// r = b - A x over the listed cells (FP32), without the norm reduction.
void residual_field(msbg_poisson *p, hipStream_t s, int storage, const void *x, const void *b, float *r);
// `sweeps` red-black sweeps in place, colour 0 then 1 (reverse: 1 then 0);
// skip_first: the first half sweep is already done.
void rb_sweeps(msbg_poisson *p, hipStream_t s, int storage, void *x, const void *b, float omega, int sweeps,
               bool reverse, bool skip_first = false);

}  // namespace pband

namespace mg {

// Builds the coarse levels of p's band (once; host tables + device fields).
int ensure(msbg_poisson *p);
void release(msbg_poisson *p);
// Levels including the band itself (0 before ensure).
int n_levels(const msbg_poisson *p);
// One V-cycle on A e = rt (p->d_e, p->d_rt in `storage`, e = 0 on entry, or
// e after the first red half sweep when first_done): nu pre- and
// post-smoothing red-black sweeps per level; p->d_r is scratch.
void vcycle(msbg_poisson *p, hipStream_t s, int storage, float omega, int nu, bool first_done);

}  // namespace mg

namespace learned {

// One learned correction of x (FP32): network input from the residual in
// p->d_r (of x, with its norms in p->d_norm / p->d_scale) and the phase field,
// delta from the network, x_try = x + alpha delta with alpha minimising
// ||r - alpha A delta||, accepted when ||b - A x_try|| < ||b - A x||.  On
// return p->d_r / d_norm / d_scale hold the residual of the (possibly
// updated) x; r2 / b2 receive its norms.
int correct(msbg_poisson *p, msbg_learned *net, const msbg_grid *phi, float *x, const float *b, double r2_in,
            double *r2, double *b2, int *accepted);

}  // namespace learned
}  // namespace msbg

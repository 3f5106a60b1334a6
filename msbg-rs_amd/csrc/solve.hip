// Poisson relaxation on the band of a grid's value blocks: red-black SOR and
// weighted Jacobi sweeps of the 7-point operator with FP32 / FP16 / BF16 block
// storage and FP32 arithmetic, the residual with FP64 norm reductions, and the
// solve loop (plain FP32, or FP32 iterative refinement around low-precision
// inner sweeps).
//
// The reference has no solver (docs/roadmap.md:610-702 "HAVEN'T STARTED"; the
// Jacobi formula is planned at docs/roadmap.md:622-626).  Operator and
// rounding are defined here:
//   (A x)_c = 6 x_c - (x_{-x} + x_{+x} + x_{-y} + x_{+y} + x_{-z} + x_{+z})
// over the cells of the listed value blocks; a neighbour outside the domain is
// the cell itself (Neumann mirror, math/boundary.rs:35-56), a neighbour in an
// absent / full block reads slot 0 / 1 of the field (the grid's dummy slots),
// a neighbour in an unlisted value block reads its stored value (Dirichlet).
// One relaxation of cell c:
//   s = ((((x_{-x} + x_{+x}) + x_{-y}) + x_{+y}) + x_{-z}) + x_{+z}
//   t = fmaf(-6, x_c, b_c + s)            (= (b - A x)_c)
//   x_c' = fmaf(omega / 6, t, x_c)        rounded to the storage type (RNE)
#include <cmath>
#include <type_traits>

#include "poisson.hpp"

namespace msbg {
namespace {

using i64 = int64_t;

int elem_bytes(int storage) { return storage == MSBG_STORAGE_F32 ? 4 : 2; }

// One row vector of W consecutive cells as stored (W = 4 for FP32, 8 for the
// 16-bit types: 16 bytes, one load per lane either way).  The 16-bit types
// stay packed in registers and are widened where a value is used, which keeps
// the register cost of a column at half its FP32 size.
template <int ST>
struct Row {
    static constexpr int W = ST == MSBG_STORAGE_F32 ? 4 : 8;
    uint32_t w[4];
    __device__ __forceinline__ float at(int i) const {
        if constexpr (ST == MSBG_STORAGE_F32)
            return __uint_as_float(w[i]);
        else
            return (i & 1) ? Sto<ST>::hi(w[i >> 1]) : Sto<ST>::lo(w[i >> 1]);
    }
    __device__ __forceinline__ void load(const typename Sto<ST>::T *p, i64 i) {  // i: multiple of W
        const uint4 q = *reinterpret_cast<const uint4 *>(p + i);
        w[0] = q.x;
        w[1] = q.y;
        w[2] = q.z;
        w[3] = q.w;
    }
    __device__ __forceinline__ void store(typename Sto<ST>::T *p, i64 i) const {
        *reinterpret_cast<uint4 *>(p + i) = make_uint4(w[0], w[1], w[2], w[3]);
    }
    // Streaming store: the row is not read again by this launch.
    __device__ __forceinline__ void store_nt(typename Sto<ST>::T *p, i64 i) const {
        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
        const u32x4 v = {w[0], w[1], w[2], w[3]};
        __builtin_nontemporal_store(v, reinterpret_cast<u32x4 *>(p + i));
    }
    __device__ __forceinline__ typename Sto<ST>::T bits(int i) const {
        if constexpr (ST == MSBG_STORAGE_F32)
            return __uint_as_float(w[i]);
        else
            return (typename Sto<ST>::T)((i & 1) ? w[i >> 1] >> 16 : w[i >> 1] & 0xffffu);
    }
    // Cell i takes v (rounded to the storage type); the other cells keep their bits.
    __device__ __forceinline__ void set(int i, float v) {
        if constexpr (ST == MSBG_STORAGE_F32) {
            w[i] = __float_as_uint(v);
        } else {
            const uint32_t h = Sto<ST>::cv(v);
            w[i >> 1] = (i & 1) ? (w[i >> 1] & 0xffffu) | h << 16 : (w[i >> 1] & 0xffff0000u) | h;
        }
    }
};

// A column: W x-adjacent cells (x = W q .. W q + W - 1) of one row, at Z
// consecutive z, of a listed block, with their 6 neighbours resolved as in the
// operator: a neighbour outside the domain is the cell itself, across a block
// face it is read from the face-neighbour slot.  A work-group (256 threads)
// covers G = 256 / (16 / W * 16) z-groups of Z slices; KS work-groups cover a
// block.  Every value a thread needs is loaded up front (one 16-byte load per
// row vector), which keeps enough bytes in flight per wave for HBM-sized bands.
template <int ST, int Z>
struct ColMap {
    static constexpr int W = Row<ST>::W;
    static constexpr int NQ = 16 / W;               // columns per row
    static constexpr int G = kThreads / (NQ * 16);  // z-groups per work-group
    static constexpr int KS = 16 / (G * Z);         // work-groups per block
    static_assert(KS >= 1 && KS * G * Z == 16, "column shape must tile a 16^3 block");
};

template <int ST, int Z>
struct Col {
    Row<ST> c[Z + 2];  // planes z0 - 1 .. z0 + Z; the cells are c[1 .. Z]
    Row<ST> ym[Z], yp[Z], b[Z];  // b: right-hand side rows of the cells
    float l[Z], r[Z];
    i64 c0;                   // element index of cell (x0, y, z0)
    int gx0, gy, gz0, n, nz;  // n / nz: cells inside the domain along x / z (0 = none)
};

// Column origin of this thread (element index c0, global cell gx0, gy, gz0,
// in-block vid); false when the column lies outside the domain.
template <int ST, int Z>
__device__ __forceinline__ bool col_origin(const PBlock &B, const Dom &d, int wg, Col<ST, Z> &Q, int &vid) {
    using M = ColMap<ST, Z>;
    const int t = threadIdx.x, q = t % M::NQ, ty = (t / M::NQ) & 15;
    const int tz0 = ((wg % M::KS) * M::G + t / (M::NQ * 16)) * Z;
    Q.gx0 = B.ox + M::W * q;
    Q.gy = B.oy + ty;
    Q.gz0 = B.oz + tz0;
    Q.n = Q.nz = 0;
    if (Q.gx0 >= d.sx || Q.gy >= d.sy || Q.gz0 >= d.sz) return false;
    Q.n = min(M::W, d.sx - Q.gx0);
    Q.nz = min(Z, d.sz - Q.gz0);
    vid = M::W * q | ty << 4 | tz0 << 8;
    Q.c0 = ((i64)B.slot << 12) + vid;
    return true;
}

// Every neighbour is a load from a selected address (a mirrored neighbour is
// the cell itself, so it is loaded from the cell's own address), so no load
// waits on another and the compiler issues them back to back.  Rows past the
// domain (j >= nz) repeat the last in-domain row; their values are unused.
template <int ST, int Z>
__device__ __forceinline__ void load_col(const PBlock &B, const Dom &d, const typename Sto<ST>::T *x,
                                         const typename Sto<ST>::T *b, int wg, Col<ST, Z> &Q) {
    using M = ColMap<ST, Z>;
    constexpr int W = M::W;
    int vid;
    if (!col_origin<ST, Z>(B, d, wg, Q, vid)) return;
    const int q = (vid & 15) / W, ty = (vid >> 4) & 15, tz0 = vid >> 8;
    const i64 nb[6] = {(i64)B.nb[0] << 12, (i64)B.nb[1] << 12, (i64)B.nb[2] << 12,
                       (i64)B.nb[3] << 12, (i64)B.nb[4] << 12, (i64)B.nb[5] << 12};
    // z - 1 plane, z + Z plane (the latter read only when the column's last
    // cell has an in-domain +z neighbour)
    Q.c[0].load(x, tz0 > 0 ? Q.c0 - 256 : (Q.gz0 == 0 ? Q.c0 : nb[4] + (vid | 0xf00)));
    Q.c[Z + 1].load(x, Q.gz0 + Z >= d.sz ? Q.c0 : (tz0 + Z <= 15 ? Q.c0 + 256 * Z : nb[5] + (vid & ~0xf00)));
#pragma unroll
    for (int j = 0; j < Z; ++j) {
        const int jj = min(j, Q.nz - 1);
        const i64 cj = Q.c0 + 256 * jj;
        const int vj = vid + 256 * jj;
        Q.c[j + 1].load(x, cj);
        Q.b[j].load(b, cj);
        Q.ym[j].load(x, ty > 0 ? cj - 16 : (Q.gy == 0 ? cj : nb[2] + (vj | 0xf0)));
        Q.yp[j].load(x, Q.gy + 1 >= d.sy ? cj : (ty < 15 ? cj + 16 : nb[3] + (vj & ~0xf0)));
        Q.l[j] = Sto<ST>::ld(x, q > 0 ? cj - 1 : (Q.gx0 == 0 ? cj : nb[0] + (vj | 15)));
        Q.r[j] = Sto<ST>::ld(x, Q.gx0 + W >= d.sx ? cj : (q < M::NQ - 1 ? cj + W : nb[1] + (vj & ~15)));
    }
}

// t = fmaf(-6, x_c, b_c + s) of cell (i, j) of the column, s summed -x, +x, -y, +y, -z, +z.
template <int ST, int Z>
__device__ __forceinline__ float col_residual(const Col<ST, Z> &Q, const Dom &d, int i, int j, float bi) {
    constexpr int W = ColMap<ST, Z>::W;
    const float xc = Q.c[j + 1].at(i);
    const float xl = i == 0 ? Q.l[j] : Q.c[j + 1].at(i - 1);
    const float xr = Q.gx0 + i + 1 >= d.sx ? xc : (i == W - 1 ? Q.r[j] : Q.c[j + 1].at(i + 1));
    const float zp = Q.gz0 + j + 1 >= d.sz ? xc : Q.c[j + 2].at(i);
    const float s = ((((xl + xr) + Q.ym[j].at(i)) + Q.yp[j].at(i)) + Q.c[j].at(i)) + zp;
    return __builtin_fmaf(-6.0f, xc, bi + s);
}

// RB: in place, cells with (gx + gy + gz) & 1 == color (their neighbours all
// have the other colour, so a launch never reads a cell it writes; the row
// vector is written back whole, the other colour's cells with their loaded
// bits).  !RB: weighted Jacobi, x_in -> x_out (distinct buffers; cells
// outside the domain are not written).
template <int ST, bool RB, int Z>
__global__ __launch_bounds__(kThreads) void relax_kernel(const PBlock *__restrict__ blocks, Dom d,
                                                         const typename Sto<ST>::T *xin, typename Sto<ST>::T *xout,
                                                         const typename Sto<ST>::T *b, float w6, int color) {
    constexpr int W = ColMap<ST, Z>::W;
    const int wg = xcd_remap(blockIdx.x, gridDim.x);
    const PBlock B = blocks[wg / ColMap<ST, Z>::KS];
    Col<ST, Z> Q;
    load_col<ST, Z>(B, d, xin, b, wg, Q);
    if (!Q.n) return;
#pragma unroll
    for (int j = 0; j < Z; ++j) {
        if (j >= Q.nz) break;
        const int par = (Q.gx0 + Q.gy + Q.gz0 + j) & 1;
        Row<ST> v = Q.c[j + 1];
#pragma unroll
        for (int i = 0; i < W; ++i)
            if (i < Q.n && (!RB || ((par + i) & 1) == color))
                v.set(i, __builtin_fmaf(w6, col_residual<ST, Z>(Q, d, i, j, Q.b[j].at(i)), Q.c[j + 1].at(i)));
        const i64 cj = Q.c0 + 256 * j;
        if (RB) {
            v.store(xout, cj);
        } else if (Q.n == W) {
            v.store_nt(xout, cj);
        } else {
#pragma unroll
            for (int i = 0; i < W; ++i)
                if (i < Q.n) xout[cj + i] = v.bits(i);
        }
    }
}

// Residual r = b - A x (r optional, FP32) and per-work-group FP64 partial sums
// of r^2 and b^2 (fixed-order reduction: per thread cells in (z, x) order,
// lane butterfly, then waves 0..3).
template <int ST, int Z>
__global__ __launch_bounds__(kThreads) void residual_kernel(const PBlock *__restrict__ blocks, Dom d,
                                                            const typename Sto<ST>::T *x,
                                                            const typename Sto<ST>::T *b, float *r,
                                                            double2 *partial) {
    constexpr int W = ColMap<ST, Z>::W;
    const int wg = xcd_remap(blockIdx.x, gridDim.x);
    const PBlock B = blocks[wg / ColMap<ST, Z>::KS];
    Col<ST, Z> Q;
    load_col<ST, Z>(B, d, x, b, wg, Q);
    double rr = 0.0, bb = 0.0;
    if (Q.n) {
#pragma unroll
        for (int j = 0; j < Z; ++j) {
            if (j >= Q.nz) break;
            float t[W];
#pragma unroll
            for (int i = 0; i < W; ++i) {
                t[i] = 0.0f;
                if (i < Q.n) {
                    const float bi = Q.b[j].at(i);
                    t[i] = col_residual<ST, Z>(Q, d, i, j, bi);
                    rr += (double)t[i] * (double)t[i];
                    bb += (double)bi * (double)bi;
                }
            }
            if (r) {
                const i64 cj = Q.c0 + 256 * j;
                if (Q.n == W) {
#pragma unroll
                    for (int k = 0; k < W; k += 4)
                        *reinterpret_cast<float4 *>(r + cj + k) = make_float4(t[k], t[k + 1], t[k + 2], t[k + 3]);
                } else {
#pragma unroll
                    for (int i = 0; i < W; ++i)
                        if (i < Q.n) r[cj + i] = t[i];
                }
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        rr += __shfl_xor(rr, o);
        bb += __shfl_xor(bb, o);
    }
    __shared__ double s_rr[kThreads / 64], s_bb[kThreads / 64];
    if ((threadIdx.x & 63) == 0) {
        s_rr[threadIdx.x >> 6] = rr;
        s_bb[threadIdx.x >> 6] = bb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, c = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) {
            a += s_rr[w];
            c += s_bb[w];
        }
        partial[blockIdx.x] = make_double2(a, c);
    }
}

// Sums the partials (thread t takes t, t + 256, ..., then a fixed tree) into
// norm[0] = ||r||^2, norm[1] = ||b||^2, and sets the refinement scale
// scale[0] = 2^e ~ rms(r) (exact power of two, 1 when r = 0), scale[1] = 1 / scale[0].
__global__ __launch_bounds__(256) void reduce_norms_kernel(const double2 *__restrict__ partial, int n, double n_cells,
                                                           double *norm, float *scale) {
    __shared__ double s_a[256], s_b[256];
    double a = 0.0, b = 0.0;
    // unrolled so a thread's next partials are in flight while it adds (the
    // same terms in the same order)
#pragma unroll 8
    for (int i = threadIdx.x; i < n; i += 256) {
        a += partial[i].x;
        b += partial[i].y;
    }
    s_a[threadIdx.x] = a;
    s_b[threadIdx.x] = b;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            s_a[threadIdx.x] += s_a[threadIdx.x + o];
            s_b[threadIdx.x] += s_b[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        norm[0] = s_a[0];
        norm[1] = s_b[0];
        int e = 0;
        const double rms = n_cells > 0 ? sqrt(s_a[0] / n_cells) : 0.0;
        if (rms > 0.0 && isfinite(rms)) {
            frexp(rms, &e);  // rms = m 2^e, m in [0.5, 1)
            e -= 1;
            e = e < -120 ? -120 : (e > 120 ? 120 : e);
        }
        scale[0] = ldexpf(1.0f, e);
        scale[1] = ldexpf(1.0f, -e);
    }
}

// One relaxation of a cell whose value and neighbours are all +0 (b: its
// right-hand side as stored), in the relaxation's operation order.
__device__ __forceinline__ float first_half(float b, float w6) {
    const float s = ((((0.0f + 0.0f) + 0.0f) + 0.0f) + 0.0f) + 0.0f;
    return __builtin_fmaf(w6, __builtin_fmaf(-6.0f, 0.0f, b + s), 0.0f);
}

// Refinement step start: rt = T(r / scale * extra), e = 0 on the listed cells
// (extra: an exact power of two).  FIRST: e also takes the first red half
// sweep from e = 0, computed as the relaxation computes it (every neighbour
// +0): colour 0 gets fmaf(w6, fmaf(-6, 0, rt + 0), 0), colour 1 stays +0.
// (column mapping of the relaxation, row vectors where the row is whole).
template <int ST, int Z, bool FIRST>
__global__ __launch_bounds__(kThreads) void ir_scale_kernel(const PBlock *__restrict__ blocks, Dom d, const float *r,
                                                            const float *scale, float extra, typename Sto<ST>::T *rt,
                                                            typename Sto<ST>::T *e, float w6) {
    constexpr int W = ColMap<ST, Z>::W;
    const int wg = xcd_remap(blockIdx.x, gridDim.x);
    const PBlock B = blocks[wg / ColMap<ST, Z>::KS];
    Col<ST, Z> Q;
    int vid;
    if (!col_origin<ST, Z>(B, d, wg, Q, vid)) return;
    const float inv = scale[1] * extra;
#pragma unroll
    for (int j = 0; j < Z; ++j) {
        if (j >= Q.nz) break;
        const i64 cj = Q.c0 + 256 * j;
        if (Q.n == W) {
            Row<ST> v, z;
#pragma unroll
            for (int k = 0; k < 4; ++k) v.w[k] = z.w[k] = 0u;  // +0 in every storage type
#pragma unroll
            for (int k = 0; k < W; k += 4) {
                const float4 q = *reinterpret_cast<const float4 *>(r + cj + k);
                v.set(k, q.x * inv);
                v.set(k + 1, q.y * inv);
                v.set(k + 2, q.z * inv);
                v.set(k + 3, q.w * inv);
            }
            if (FIRST) {
                const int par = (Q.gx0 + Q.gy + Q.gz0 + j) & 1;
#pragma unroll
                for (int i = 0; i < W; ++i)
                    if (((par + i) & 1) == 0) z.set(i, first_half(v.at(i), w6));
            }
            v.store(rt, cj);
            z.store(e, cj);
        } else {
            const int par = (Q.gx0 + Q.gy + Q.gz0 + j) & 1;
            for (int i = 0; i < Q.n; ++i) {
                const typename Sto<ST>::T t = Sto<ST>::cv(r[cj + i] * inv);
                rt[cj + i] = t;
                e[cj + i] = Sto<ST>::cv(FIRST && ((par + i) & 1) == 0 ? first_half(Sto<ST>::ld(&t, 0), w6) : 0.0f);
            }
        }
    }
}

// Refinement step end: x = fmaf(scale / extra, e, x) on the listed cells.
template <int ST, int Z>
__global__ __launch_bounds__(kThreads) void ir_update_kernel(const PBlock *__restrict__ blocks, Dom d,
                                                             const typename Sto<ST>::T *e, const float *scale,
                                                             float extra, float *x) {
    constexpr int W = ColMap<ST, Z>::W;
    const int wg = xcd_remap(blockIdx.x, gridDim.x);
    const PBlock B = blocks[wg / ColMap<ST, Z>::KS];
    Col<ST, Z> Q;
    int vid;
    if (!col_origin<ST, Z>(B, d, wg, Q, vid)) return;
    const float s = scale[0] / extra;
#pragma unroll
    for (int j = 0; j < Z; ++j) {
        if (j >= Q.nz) break;
        const i64 cj = Q.c0 + 256 * j;
        if (Q.n == W) {
            Row<ST> ev;
            ev.load(e, cj);
#pragma unroll
            for (int k = 0; k < W; k += 4) {
                float4 q = *reinterpret_cast<const float4 *>(x + cj + k);
                q.x = __builtin_fmaf(s, ev.at(k), q.x);
                q.y = __builtin_fmaf(s, ev.at(k + 1), q.y);
                q.z = __builtin_fmaf(s, ev.at(k + 2), q.z);
                q.w = __builtin_fmaf(s, ev.at(k + 3), q.w);
                *reinterpret_cast<float4 *>(x + cj + k) = q;
            }
        } else {
            for (int i = 0; i < Q.n; ++i) x[cj + i] = __builtin_fmaf(s, Sto<ST>::ld(e, cj + i), x[cj + i]);
        }
    }
}

// ---------------------------------------------------------------------------
// Compensated iterate: x carried as the unevaluated sum hi + lo of two FP32
// fields (|lo| <= ulp(hi) / 2), so the refinement can take the residual below
// the rounding floor of one FP32 field (~2.3e-7 ||x|| / ||b||).  Error-free
// transformations (Knuth TwoSum, Dekker FastTwoSum, fma TwoProd) in FP32.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float two_sum(float a, float b, float &err) {
    const float s = a + b;
    const float bb = s - a;
    err = (a - (s - bb)) + (b - bb);
    return s;
}

// (hi, lo) + v, renormalised.
__device__ __forceinline__ void pair_add(float &hi, float &lo, float v) {
    float t;
    const float h = two_sum(hi, v, t);
    float u;
    hi = two_sum(h, lo + t, u);
    lo = u;
}

// r = b - A (hi + lo) of column cell (i, j), rounded once to FP32: the hi part
// of A x (6 hi_c and the six hi neighbours) and b are summed error-free, the
// lo part (|A lo| ~ 2^-22 |x|) in plain FP32.
template <int Z>
__device__ __forceinline__ float col_residual_pair(const Col<MSBG_STORAGE_F32, Z> &Q,
                                                   const Col<MSBG_STORAGE_F32, Z> &QL, const Dom &d, int i, int j,
                                                   float bi) {
    constexpr int W = 4;
    auto taps = [&](const Col<MSBG_STORAGE_F32, Z> &C, float *t) {
        const float xc = C.c[j + 1].at(i);
        t[0] = xc;
        t[1] = i == 0 ? C.l[j] : C.c[j + 1].at(i - 1);
        t[2] = Q.gx0 + i + 1 >= d.sx ? xc : (i == W - 1 ? C.r[j] : C.c[j + 1].at(i + 1));
        t[3] = C.ym[j].at(i);
        t[4] = C.yp[j].at(i);
        t[5] = C.c[j].at(i);
        t[6] = Q.gz0 + j + 1 >= d.sz ? xc : C.c[j + 2].at(i);
    };
    float h[7], l[7];
    taps(Q, h);
    taps(QL, l);
    const float p6 = 6.0f * h[0];
    float err = __builtin_fmaf(6.0f, h[0], -p6);  // 6 hi_c = p6 + err exactly
    float e, acc = two_sum(bi, -p6, e);
    e -= err;
#pragma unroll
    for (int k = 1; k < 7; ++k) {
        float t;
        acc = two_sum(acc, h[k], t);
        e += t;
    }
    const float slo = ((((l[1] + l[2]) + l[3]) + l[4]) + l[5]) + l[6];
    e += __builtin_fmaf(-6.0f, l[0], slo);
    return acc + e;
}

// Residual of a compensated FP32 iterate, with the FP64 partial norms of
// residual_kernel (same reduction order).
template <int Z>
__global__ __launch_bounds__(kThreads) void residual_pair_kernel(const PBlock *__restrict__ blocks, Dom d,
                                                                 const float *x, const float *xlo, const float *b,
                                                                 float *r, double2 *partial) {
    constexpr int W = 4;
    const int wg = xcd_remap(blockIdx.x, gridDim.x);
    const PBlock B = blocks[wg / ColMap<MSBG_STORAGE_F32, Z>::KS];
    Col<MSBG_STORAGE_F32, Z> Q, QL;
    load_col<MSBG_STORAGE_F32, Z>(B, d, x, b, wg, Q);
    load_col<MSBG_STORAGE_F32, Z>(B, d, xlo, xlo, wg, QL);
    double rr = 0.0, bb = 0.0;
    if (Q.n) {
#pragma unroll
        for (int j = 0; j < Z; ++j) {
            if (j >= Q.nz) break;
            float t[W];
#pragma unroll
            for (int i = 0; i < W; ++i) {
                t[i] = 0.0f;
                if (i < Q.n) {
                    const float bi = Q.b[j].at(i);
                    t[i] = col_residual_pair<Z>(Q, QL, d, i, j, bi);
                    rr += (double)t[i] * (double)t[i];
                    bb += (double)bi * (double)bi;
                }
            }
            if (r) {
                const i64 cj = Q.c0 + 256 * j;
                if (Q.n == W) {
                    *reinterpret_cast<float4 *>(r + cj) = make_float4(t[0], t[1], t[2], t[3]);
                } else {
#pragma unroll
                    for (int i = 0; i < W; ++i)
                        if (i < Q.n) r[cj + i] = t[i];
                }
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        rr += __shfl_xor(rr, o);
        bb += __shfl_xor(bb, o);
    }
    __shared__ double s_rr[kThreads / 64], s_bb[kThreads / 64];
    if ((threadIdx.x & 63) == 0) {
        s_rr[threadIdx.x >> 6] = rr;
        s_bb[threadIdx.x >> 6] = bb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, c = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) {
            a += s_rr[w];
            c += s_bb[w];
        }
        partial[blockIdx.x] = make_double2(a, c);
    }
}

// Refinement step end on a compensated iterate: (x, lo) += (scale / extra) e
// (the product is exact: both factors of the scale are powers of two).
template <int ST, int Z>
__global__ __launch_bounds__(kThreads) void ir_update_pair_kernel(const PBlock *__restrict__ blocks, Dom d,
                                                                  const typename Sto<ST>::T *e, const float *scale,
                                                                  float extra, float *x, float *xlo) {
    constexpr int W = ColMap<ST, Z>::W;
    const int wg = xcd_remap(blockIdx.x, gridDim.x);
    const PBlock B = blocks[wg / ColMap<ST, Z>::KS];
    Col<ST, Z> Q;
    int vid;
    if (!col_origin<ST, Z>(B, d, wg, Q, vid)) return;
    const float s = scale[0] / extra;
#pragma unroll
    for (int j = 0; j < Z; ++j) {
        if (j >= Q.nz) break;
        const i64 cj = Q.c0 + 256 * j;
        if (Q.n == W) {  // whole row vectors: one 16-byte load of e, float4 loads and stores of hi and lo
            Row<ST> ev;
            ev.load(e, cj);
#pragma unroll
            for (int k = 0; k < W; k += 4) {
                float4 h = *reinterpret_cast<const float4 *>(x + cj + k);
                float4 l = *reinterpret_cast<const float4 *>(xlo + cj + k);
                pair_add(h.x, l.x, s * ev.at(k));
                pair_add(h.y, l.y, s * ev.at(k + 1));
                pair_add(h.z, l.z, s * ev.at(k + 2));
                pair_add(h.w, l.w, s * ev.at(k + 3));
                *reinterpret_cast<float4 *>(x + cj + k) = h;
                *reinterpret_cast<float4 *>(xlo + cj + k) = l;
            }
        } else {
            for (int i = 0; i < Q.n; ++i) {
                float hi = x[cj + i], lo = xlo[cj + i];
                pair_add(hi, lo, s * Sto<ST>::ld(e, cj + i));
                x[cj + i] = hi;
                xlo[cj + i] = lo;
            }
        }
    }
}

__global__ void zero_u32_kernel(uint32_t *p, i64 n) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) p[i] = 0u;
}

// Storage conversion of whole slabs (RNE on narrowing).
template <int SS, int DS>
__global__ void convert_kernel(const typename Sto<SS>::T *src, typename Sto<DS>::T *dst, i64 n) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        dst[i] = Sto<DS>::cv(Sto<SS>::ld(src, i));
}

// Column depth Z of the relaxation / residual / refinement kernels of storage
// `st` (p->zcol), as a compile-time constant for f.
template <int ST, class F>
void with_z(const msbg_poisson *p, F &&f) {
    const int z = p->zcol[ST];
    if constexpr (ST == MSBG_STORAGE_F32) {
        if (z == 4)
            f(std::integral_constant<int, 4>{});
        else if (z == 2)
            f(std::integral_constant<int, 2>{});
        else
            f(std::integral_constant<int, 1>{});
    } else {
        if (z == 2)
            f(std::integral_constant<int, 2>{});
        else
            f(std::integral_constant<int, 1>{});
    }
}

template <int ST, int Z>
int col_groups(const msbg_poisson *p) {
    return (int)(p->n_active * ColMap<ST, Z>::KS);
}

}  // namespace
}  // namespace msbg

using namespace msbg;

namespace msbg {
namespace pband {

template <int ST>
void residual_st(msbg_poisson *p, hipStream_t s, const void *x, const void *b, float *r) {
    using T = typename Sto<ST>::T;
    with_z<ST>(p, [&](auto zc) {
        constexpr int Z = decltype(zc)::value;
        const int n = col_groups<ST, Z>(p);
        residual_kernel<ST, Z><<<n, kThreads, 0, s>>>(p->d_blocks, p->dom, (const T *)x, (const T *)b, r,
                                                      p->d_partial);
        reduce_norms_kernel<<<1, 256, 0, s>>>(p->d_partial, n, (double)p->n_cells, p->d_norm, p->d_scale);
    });
}

void residual(msbg_poisson *p, hipStream_t s, int st, const void *x, const void *b, float *r) {
    if (st == MSBG_STORAGE_F32)
        residual_st<MSBG_STORAGE_F32>(p, s, x, b, r);
    else if (st == MSBG_STORAGE_F16)
        residual_st<MSBG_STORAGE_F16>(p, s, x, b, r);
    else
        residual_st<MSBG_STORAGE_BF16>(p, s, x, b, r);
}

int read_norms(msbg_poisson *p, double *r2, double *b2) {
    MSBG_HIP(hipMemcpyAsync(p->h_norm, p->d_norm, 2 * sizeof(double), hipMemcpyDeviceToHost, stream()));
    MSBG_HIP(hipStreamSynchronize(stream()));
    *r2 = p->h_norm[0];
    *b2 = p->h_norm[1];
    return MSBG_OK;
}

}  // namespace pband

namespace {

using pband::read_norms;
using pband::residual;

void drop_graph(msbg_poisson *p) {
    if (p->exec) {
        if (p->g_done) (void)hipEventSynchronize(p->g_done);
        (void)hipGraphExecDestroy(p->exec);
    }
    p->exec = nullptr;
    p->g_mode = -1;
}

void drop_jacobi_graph(msbg_poisson *p) {
    if (p->jexec) {
        if (p->j_done) (void)hipEventSynchronize(p->j_done);
        (void)hipGraphExecDestroy(p->jexec);
    }
    p->jexec = nullptr;
    p->j_sweeps = -1;
}

int check_storage(int st) {
    MSBG_CHECK(st == MSBG_STORAGE_F32 || st == MSBG_STORAGE_F16 || st == MSBG_STORAGE_BF16, MSBG_EINVAL,
               "storage must be MSBG_STORAGE_F32, _F16 or _BF16");
    return MSBG_OK;
}

template <int ST>
void launch_relax(msbg_poisson *p, hipStream_t s, const void *xin, void *xout, const void *b, float omega, int color,
                  bool rb) {
    using T = typename Sto<ST>::T;
    const float w6 = omega / 6.0f;
    with_z<ST>(p, [&](auto zc) {
        constexpr int Z = decltype(zc)::value;
        const int n = col_groups<ST, Z>(p);
        if (rb)
            relax_kernel<ST, true, Z><<<n, kThreads, 0, s>>>(p->d_blocks, p->dom, (const T *)xin, (T *)xout,
                                                             (const T *)b, w6, color);
        else
            relax_kernel<ST, false, Z><<<n, kThreads, 0, s>>>(p->d_blocks, p->dom, (const T *)xin, (T *)xout,
                                                              (const T *)b, w6, color);
    });
}

void relax(msbg_poisson *p, hipStream_t s, int st, const void *xin, void *xout, const void *b, float omega, int color,
           bool rb) {
    if (st == MSBG_STORAGE_F32)
        launch_relax<MSBG_STORAGE_F32>(p, s, xin, xout, b, omega, color, rb);
    else if (st == MSBG_STORAGE_F16)
        launch_relax<MSBG_STORAGE_F16>(p, s, xin, xout, b, omega, color, rb);
    else
        launch_relax<MSBG_STORAGE_BF16>(p, s, xin, xout, b, omega, color, rb);
}

void rb_sweeps(msbg_poisson *p, hipStream_t s, int st, void *x, const void *b, float omega, int sweeps) {
    pband::rb_sweeps(p, s, st, x, b, omega, sweeps, false);
}

template <int ST>
void ir_scale_st(msbg_poisson *p, hipStream_t s, float extra, float w6) {
    using T = typename Sto<ST>::T;
    with_z<ST>(p, [&](auto zc) {
        constexpr int Z = decltype(zc)::value;
        if (w6 > 0.0f)
            ir_scale_kernel<ST, Z, true><<<col_groups<ST, Z>(p), kThreads, 0, s>>>(
                p->d_blocks, p->dom, p->d_r, p->d_scale, extra, (T *)p->d_rt, (T *)p->d_e, w6);
        else
            ir_scale_kernel<ST, Z, false><<<col_groups<ST, Z>(p), kThreads, 0, s>>>(
                p->d_blocks, p->dom, p->d_r, p->d_scale, extra, (T *)p->d_rt, (T *)p->d_e, 0.0f);
    });
}

template <int ST>
void ir_update_st(msbg_poisson *p, hipStream_t s, float *x, float extra) {
    using T = typename Sto<ST>::T;
    with_z<ST>(p, [&](auto zc) {
        constexpr int Z = decltype(zc)::value;
        ir_update_kernel<ST, Z><<<col_groups<ST, Z>(p), kThreads, 0, s>>>(p->d_blocks, p->dom, (const T *)p->d_e,
                                                                           p->d_scale, extra, x);
    });
}

// w6 > 0: the first red half sweep (factor w6 = omega / 6) is folded in.
void ir_scale(msbg_poisson *p, hipStream_t s, int st, float extra, float w6) {
    if (st == MSBG_STORAGE_F32)
        ir_scale_st<MSBG_STORAGE_F32>(p, s, extra, w6);
    else if (st == MSBG_STORAGE_F16)
        ir_scale_st<MSBG_STORAGE_F16>(p, s, extra, w6);
    else
        ir_scale_st<MSBG_STORAGE_BF16>(p, s, extra, w6);
}

void ir_update(msbg_poisson *p, hipStream_t s, int st, float *x, float extra) {
    if (st == MSBG_STORAGE_F32)
        ir_update_st<MSBG_STORAGE_F32>(p, s, x, extra);
    else if (st == MSBG_STORAGE_F16)
        ir_update_st<MSBG_STORAGE_F16>(p, s, x, extra);
    else
        ir_update_st<MSBG_STORAGE_BF16>(p, s, x, extra);
}

// r = b - A (x + lo) and its norms (compensated iterate; FP32 fields).
void residual_pair(msbg_poisson *p, hipStream_t s, const float *x, const float *lo, const float *b, float *r) {
    constexpr int Z = 1;  // two columns of loads per thread: the shallowest column
    const int n = col_groups<MSBG_STORAGE_F32, Z>(p);
    residual_pair_kernel<Z><<<n, kThreads, 0, s>>>(p->d_blocks, p->dom, x, lo, b, r, p->d_partial);
    reduce_norms_kernel<<<1, 256, 0, s>>>(p->d_partial, n, (double)p->n_cells, p->d_norm, p->d_scale);
}

template <int ST>
void ir_update_pair_st(msbg_poisson *p, hipStream_t s, float *x, float *lo, float extra) {
    using T = typename Sto<ST>::T;
    with_z<ST>(p, [&](auto zc) {
        constexpr int Z = decltype(zc)::value;
        ir_update_pair_kernel<ST, Z><<<col_groups<ST, Z>(p), kThreads, 0, s>>>(p->d_blocks, p->dom,
                                                                                (const T *)p->d_e, p->d_scale,
                                                                                extra, x, lo);
    });
}

void ir_update_pair(msbg_poisson *p, hipStream_t s, int st, float *x, float *lo, float extra) {
    if (st == MSBG_STORAGE_F32)
        ir_update_pair_st<MSBG_STORAGE_F32>(p, s, x, lo, extra);
    else if (st == MSBG_STORAGE_F16)
        ir_update_pair_st<MSBG_STORAGE_F16>(p, s, x, lo, extra);
    else
        ir_update_pair_st<MSBG_STORAGE_BF16>(p, s, x, lo, extra);
}

// Residual of the iterate: plain FP32 x, or the compensated x + lo.
void residual_of(msbg_poisson *p, hipStream_t s, const float *x, const float *lo, const float *b, float *r) {
    if (lo)
        residual_pair(p, s, x, lo, b, r);
    else
        residual(p, s, MSBG_STORAGE_F32, x, b, r);
}

// Inner fields for `storage`; slots outside the listed blocks stay zero (the
// error equation has homogeneous boundary values).
int ensure_inner(msbg_poisson *p, int storage) {
    const int eb = elem_bytes(storage);
    const size_t n = (size_t)p->n_slots * 4096;
    if (!p->d_r) MSBG_HIP(hipMalloc(&p->d_r, n * 4));
    if (p->inner_bytes < eb) {
        drop_graph(p);
        if (p->d_e) (void)hipFree(p->d_e);
        if (p->d_rt) (void)hipFree(p->d_rt);
        p->d_e = p->d_rt = nullptr;
        p->inner_bytes = 0;
        MSBG_HIP(hipMalloc(&p->d_e, n * eb));
        MSBG_HIP(hipMalloc(&p->d_rt, n * eb));
        p->inner_bytes = eb;
        p->e_zeroed = false;
    }
    if (!p->e_zeroed) {
        const i64 words = (i64)(n * p->inner_bytes / 4);
        zero_u32_kernel<<<1024, 256, 0, stream()>>>((uint32_t *)p->d_e, words);
        zero_u32_kernel<<<1024, 256, 0, stream()>>>((uint32_t *)p->d_rt, words);
        MSBG_HIP(hipGetLastError());
        p->e_zeroed = true;
    }
    return MSBG_OK;
}

// One check period.  mode 0: `sweeps` red-black sweeps on x (FP32), then the
// residual.  mode 1: refinement step rt = r / s, e = 0, `sweeps` inner sweeps
// on e in `storage`, x += s e, then the residual (into r).  mode 2: the same
// refinement step with one multigrid V-cycle (`sweeps` pre- and
// post-smoothing sweeps per level) as the inner solve.
void period(msbg_poisson *p, hipStream_t s, int mode, int storage, float *x, float *lo, const float *b, float omega,
            int sweeps) {
    if (mode == 0) {
        rb_sweeps(p, s, MSBG_STORAGE_F32, x, b, omega, sweeps);
        residual(p, s, MSBG_STORAGE_F32, x, b, p->d_r);
    } else {
        // A V-cycle returns the whole correction, |e| up to ~1 / lambda_min
        // times rms(rt): the FP16 inner problem is scaled down by 2^-12 more,
        // leaving 2^12 * 65504 of headroom (entries below rms / 4 become
        // subnormal, an absolute error <= 2^-13 rms).
        const float extra = (mode == 2 && storage == MSBG_STORAGE_F16) ? 0x1p-12f : 1.0f;
        // the first red half sweep from e = 0 is computed by the scale kernel
        const bool fold = sweeps > 0;
        ir_scale(p, s, storage, extra, fold ? omega / 6.0f : 0.0f);
        if (mode == 2)
            mg::vcycle(p, s, storage, omega, sweeps, fold);
        else
            pband::rb_sweeps(p, s, storage, p->d_e, p->d_rt, omega, sweeps, false, fold);
        if (lo)
            ir_update_pair(p, s, storage, x, lo, extra);
        else
            ir_update(p, s, storage, x, extra);
        residual_of(p, s, x, lo, b, p->d_r);
    }
}

// Bands from this many cells up run their V-cycles as direct launches.
constexpr int64_t kDirectCycleCells = (int64_t)64 << 20;

int run_period_graph(msbg_poisson *p, int mode, int storage, float *x, float *lo, const float *b, float omega,
                     int sweeps) {
    if (!(p->exec && p->g_mode == mode && p->g_storage == storage && p->g_sweeps == sweeps && p->g_omega == omega &&
          p->g_x == x && p->g_lo == lo && p->g_b == b)) {
        drop_graph(p);
        if (!p->g_done) MSBG_HIP(hipEventCreateWithFlags(&p->g_done, hipEventDisableTiming));
        if (!p->cap_stream) MSBG_HIP(hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking));
        MSBG_HIP(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeRelaxed));
        period(p, p->cap_stream, mode, storage, x, lo, b, omega, sweeps);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(p->cap_stream, &graph);
        MSBG_HIP(e);
        e = hipGraphInstantiate(&p->exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) p->exec = nullptr;
        MSBG_HIP(e);
        p->g_mode = mode;
        p->g_storage = storage;
        p->g_sweeps = sweeps;
        p->g_omega = omega;
        p->g_x = x;
        p->g_lo = lo;
        p->g_b = b;
    }
    MSBG_HIP(hipGraphLaunch(p->exec, stream()));
    MSBG_HIP(hipEventRecord(p->g_done, stream()));
    return MSBG_OK;
}

}  // namespace
}  // namespace msbg

namespace msbg {
namespace pband {

template <int ST>
void residual_field_st(msbg_poisson *p, hipStream_t s, const void *x, const void *b, float *r) {
    using T = typename Sto<ST>::T;
    with_z<ST>(p, [&](auto zc) {
        constexpr int Z = decltype(zc)::value;
        residual_kernel<ST, Z><<<col_groups<ST, Z>(p), kThreads, 0, s>>>(p->d_blocks, p->dom, (const T *)x,
                                                                          (const T *)b, r, p->d_partial);
    });
}

void residual_field(msbg_poisson *p, hipStream_t s, int st, const void *x, const void *b, float *r) {
    if (st == MSBG_STORAGE_F32)
        residual_field_st<MSBG_STORAGE_F32>(p, s, x, b, r);
    else if (st == MSBG_STORAGE_F16)
        residual_field_st<MSBG_STORAGE_F16>(p, s, x, b, r);
    else
        residual_field_st<MSBG_STORAGE_BF16>(p, s, x, b, r);
}

void rb_sweeps(msbg_poisson *p, hipStream_t s, int st, void *x, const void *b, float omega, int sweeps,
               bool reverse, bool skip_first) {
    for (int i = 0; i < sweeps; ++i) {
        if (!(skip_first && i == 0)) relax(p, s, st, x, x, b, omega, reverse ? 1 : 0, true);
        relax(p, s, st, x, x, b, omega, reverse ? 0 : 1, true);
    }
}

}  // namespace pband
}  // namespace msbg

extern "C" {

int msbg_poisson_new(const msbg_grid *layout, const int64_t *active, int64_t n_active, msbg_poisson **out) {
    MSBG_CHECK(layout && out, MSBG_EINVAL, "null argument");
    MSBG_CHECK(layout->L == 4, MSBG_EINVAL, "the Poisson band needs 16^3 blocks");
    MSBG_CHECK(n_active >= 0 && (n_active == 0 || active), MSBG_EINVAL, "bad active list");
    MSBG_CHECK(n_active <= (int64_t)1 << 28, MSBG_EINVAL, "too many blocks");
    *out = nullptr;
    std::vector<PBlock> host((size_t)n_active);
    std::vector<char> seen((size_t)layout->n_blocks, 0);
    for (int64_t k = 0; k < n_active; ++k) {
        const int64_t bid = active[k];
        MSBG_CHECK(bid >= 0 && bid < layout->n_blocks, MSBG_EINVAL, "block id out of range");
        MSBG_CHECK(!seen[bid], MSBG_EINVAL, "duplicate block id in the active list");
        seen[bid] = 1;
        const int32_t slot = layout->b2s_host[bid];
        MSBG_CHECK(slot >= 2, MSBG_EINVAL, "listed block is not a value block");
        const int64_t bx = bid % layout->nx, by = (bid / layout->nx) % layout->ny, bz = bid / layout->nxy;
        PBlock &B = host[k];
        B.slot = slot;
        B.ox = (int32_t)(bx * 16);
        B.oy = (int32_t)(by * 16);
        B.oz = (int32_t)(bz * 16);
        const int64_t c[3] = {bx, by, bz}, n[3] = {layout->nx, layout->ny, layout->nz};
        for (int f = 0; f < 6; ++f) {
            const int a = f >> 1, dir = (f & 1) ? 1 : -1;
            int64_t q[3] = {c[0], c[1], c[2]};
            q[a] += dir;
            B.nb[f] = 0;
            if (q[a] >= 0 && q[a] < n[a]) B.nb[f] = layout->b2s_host[q[0] + q[1] * layout->nx + q[2] * layout->nxy];
        }
    }
    msbg_poisson *p = new msbg_poisson();
    p->bids.assign(active, active + n_active);
    p->dom = Dom{(int32_t)layout->sx, (int32_t)layout->sy, (int32_t)layout->sz};
    p->n_slots = layout->n_slots;
    p->n_active = n_active;
    int64_t cells = 0;
    for (const PBlock &B : host) {
        const int64_t ex = std::min<int64_t>(16, layout->sx - B.ox), ey = std::min<int64_t>(16, layout->sy - B.oy),
                      ez = std::min<int64_t>(16, layout->sz - B.oz);
        cells += ex * ey * ez;
    }
    p->n_cells = cells;
    p->n_wg = (int)(n_active * kSlices);
    // Column depth: the deepest column (fewest work-groups, most bytes in
    // flight per wave) that still leaves >= 2048 work-groups per launch.
    p->zcol[MSBG_STORAGE_F32] = n_active >= 2048 ? 4 : (n_active >= 1024 ? 2 : 1);
    p->zcol[MSBG_STORAGE_F16] = p->zcol[MSBG_STORAGE_BF16] = n_active >= 2048 ? 2 : 1;
    p->nx = layout->nx;
    p->ny = layout->ny;
    p->nz = layout->nz;
    p->n_blocks = layout->n_blocks;
    std::vector<uint8_t> listed((size_t)layout->n_blocks, 0);
    for (int64_t k = 0; k < n_active; ++k) listed[active[k]] = 1;
    hipError_t e = hipSuccess;
    if (n_active) e = hipMalloc(&p->d_blocks, sizeof(PBlock) * n_active);
    if (e == hipSuccess && n_active)
        e = hipMemcpy(p->d_blocks, host.data(), sizeof(PBlock) * n_active, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&p->d_b2s, sizeof(int32_t) * layout->n_blocks);
    if (e == hipSuccess)
        e = hipMemcpy(p->d_b2s, layout->b2s_host.data(), sizeof(int32_t) * layout->n_blocks, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&p->d_listed, layout->n_blocks);
    if (e == hipSuccess) e = hipMemcpy(p->d_listed, listed.data(), layout->n_blocks, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&p->d_partial, sizeof(double2) * std::max(1, p->n_wg));
    if (e == hipSuccess) e = hipMalloc(&p->d_norm, 2 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&p->d_scale, 2 * sizeof(float));
    if (e == hipSuccess) e = hipHostMalloc(&p->h_norm, 2 * sizeof(double));
    if (e != hipSuccess) {
        msbg_poisson_free(p);
        return hip_fail(e, "msbg_poisson_new");
    }
    *out = p;
    return MSBG_OK;
}

int msbg_poisson_free(msbg_poisson *p) {
    if (!p) return MSBG_OK;
    drop_graph(p);
    drop_jacobi_graph(p);
    (void)hipStreamSynchronize(stream());  // the band's buffers may still be in use by queued work
    mg::release(p);
    if (p->g_done) (void)hipEventDestroy(p->g_done);
    if (p->j_done) (void)hipEventDestroy(p->j_done);
    if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
    (void)hipFree(p->d_blocks);
    (void)hipFree(p->d_partial);
    (void)hipFree(p->d_norm);
    (void)hipFree(p->d_scale);
    if (p->h_norm) (void)hipHostFree(p->h_norm);
    (void)hipFree(p->d_r);
    (void)hipFree(p->d_e);
    (void)hipFree(p->d_rt);
    (void)hipFree(p->d_b2s);
    (void)hipFree(p->d_listed);
    (void)hipFree(p->d_delta);
    (void)hipFree(p->d_xtry);
    (void)hipFree(p->d_dots);
    delete p;
    return MSBG_OK;
}

int msbg_poisson_info(const msbg_poisson *p, int64_t *n_active, int64_t *n_cells, int64_t *n_slots) {
    MSBG_CHECK(p && n_active && n_cells && n_slots, MSBG_EINVAL, "null argument");
    *n_active = p->n_active;
    *n_cells = p->n_cells;
    *n_slots = p->n_slots;
    return MSBG_OK;
}

int msbg_poisson_jacobi(msbg_poisson *p, int storage, const void *x_in, void *x_out, const void *b, float omega) {
    MSBG_CHECK(p && x_in && x_out && b, MSBG_EINVAL, "null argument");
    MSBG_CHECK(x_in != x_out, MSBG_EINVAL, "Jacobi needs distinct input and output fields");
    int rc = check_storage(storage);
    if (rc) return rc;
    if (!p->n_active) return MSBG_OK;
    relax(p, stream(), storage, x_in, x_out, b, omega, 0, false);
    MSBG_HIP(hipGetLastError());
    return MSBG_OK;
}

int msbg_poisson_jacobi_sweeps(msbg_poisson *p, int storage, void *x, void *tmp, const void *b, float omega,
                               int sweeps) {
    MSBG_CHECK(p && x && tmp && b, MSBG_EINVAL, "null argument");
    MSBG_CHECK(x != tmp, MSBG_EINVAL, "Jacobi needs distinct fields");
    MSBG_CHECK(sweeps >= 0 && sweeps % 2 == 0, MSBG_EINVAL, "sweeps must be even (the result returns to x)");
    int rc = check_storage(storage);
    if (rc) return rc;
    if (!p->n_active || !sweeps) return MSBG_OK;
    if (!(p->jexec && p->j_storage == storage && p->j_sweeps == sweeps && p->j_omega == omega && p->j_x == x &&
          p->j_t == tmp && p->j_b == b)) {
        drop_jacobi_graph(p);
        if (!p->j_done) MSBG_HIP(hipEventCreateWithFlags(&p->j_done, hipEventDisableTiming));
        if (!p->cap_stream) MSBG_HIP(hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking));
        MSBG_HIP(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeRelaxed));
        for (int i = 0; i < sweeps; i += 2) {
            relax(p, p->cap_stream, storage, x, tmp, b, omega, 0, false);
            relax(p, p->cap_stream, storage, tmp, x, b, omega, 0, false);
        }
        hipGraph_t graph = nullptr;
        MSBG_HIP(hipStreamEndCapture(p->cap_stream, &graph));
        hipError_t e = hipGraphInstantiate(&p->jexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) p->jexec = nullptr;
        MSBG_HIP(e);
        p->j_storage = storage;
        p->j_sweeps = sweeps;
        p->j_omega = omega;
        p->j_x = x;
        p->j_t = tmp;
        p->j_b = b;
    }
    MSBG_HIP(hipGraphLaunch(p->jexec, stream()));
    MSBG_HIP(hipEventRecord(p->j_done, stream()));
    return MSBG_OK;
}

int msbg_poisson_rbgs(msbg_poisson *p, int storage, void *x, const void *b, int color, float omega) {
    MSBG_CHECK(p && x && b, MSBG_EINVAL, "null argument");
    MSBG_CHECK(color == 0 || color == 1, MSBG_EINVAL, "color must be 0 or 1");
    int rc = check_storage(storage);
    if (rc) return rc;
    if (!p->n_active) return MSBG_OK;
    relax(p, stream(), storage, x, x, b, omega, color, true);
    MSBG_HIP(hipGetLastError());
    return MSBG_OK;
}

int msbg_poisson_residual(msbg_poisson *p, int storage, const void *x, const void *b, float *r, double *r2,
                          double *b2) {
    MSBG_CHECK(p && x && b && r2 && b2, MSBG_EINVAL, "null argument");
    int rc = check_storage(storage);
    if (rc) return rc;
    if (!p->n_active) {
        *r2 = *b2 = 0.0;
        return MSBG_OK;
    }
    residual(p, stream(), storage, x, b, r);
    MSBG_HIP(hipGetLastError());
    return read_norms(p, r2, b2);
}

int msbg_poisson_residual_pair(msbg_poisson *p, const float *x, const float *x_lo, const float *b, float *r,
                               double *r2, double *b2) {
    MSBG_CHECK(p && x && x_lo && b && r2 && b2, MSBG_EINVAL, "null argument");
    if (!p->n_active) {
        *r2 = *b2 = 0.0;
        return MSBG_OK;
    }
    residual_pair(p, stream(), x, x_lo, b, r);
    MSBG_HIP(hipGetLastError());
    return read_norms(p, r2, b2);
}

int msbg_poisson_convert(int src_storage, const void *src, int dst_storage, void *dst, int64_t n_elems) {
    MSBG_CHECK(src && dst && n_elems >= 0, MSBG_EINVAL, "bad argument");
    int rc = check_storage(src_storage);
    if (rc) return rc;
    rc = check_storage(dst_storage);
    if (rc) return rc;
    if (!n_elems) return MSBG_OK;
    const int grid = (int)std::min<int64_t>(4096, (n_elems + 255) / 256);
    hipStream_t s = stream();
#define MSBG_CONV(A, B)                                                                                          \
    if (src_storage == A && dst_storage == B)                                                                    \
        convert_kernel<A, B><<<grid, 256, 0, s>>>((const typename Sto<A>::T *)src, (typename Sto<B>::T *)dst, \
                                                  n_elems);
    MSBG_CONV(0, 0) MSBG_CONV(0, 1) MSBG_CONV(0, 2) MSBG_CONV(1, 0) MSBG_CONV(1, 1) MSBG_CONV(1, 2)
    MSBG_CONV(2, 0) MSBG_CONV(2, 1) MSBG_CONV(2, 2)
#undef MSBG_CONV
    MSBG_HIP(hipGetLastError());
    return MSBG_OK;
}

int msbg_poisson_solve(msbg_poisson *p, const msbg_poisson_params *prm, float *x, const float *b,
                       msbg_poisson_stats *st) {
    MSBG_CHECK(p && prm && x && b && st, MSBG_EINVAL, "null argument");
    int rc = check_storage(prm->storage);
    if (rc) return rc;
    MSBG_CHECK(prm->refine >= 0 && prm->refine <= 2, MSBG_EINVAL, "refine must be 0, 1 or 2");
    MSBG_CHECK(prm->refine || prm->storage == MSBG_STORAGE_F32, MSBG_EINVAL,
               "sweeps on x itself need FP32 storage (use refine = 1 for FP16 / BF16)");
    MSBG_CHECK(prm->omega > 0.0f && prm->omega < 2.0f, MSBG_EINVAL, "omega must be in (0, 2)");
    MSBG_CHECK(prm->sweeps_per_check >= 1 && prm->max_sweeps >= 0, MSBG_EINVAL, "bad sweep counts");
    MSBG_CHECK(prm->tol >= 0.0, MSBG_EINVAL, "negative tolerance");
    MSBG_CHECK((prm->flags & ~MSBG_SOLVE_DIRECT) == 0, MSBG_EINVAL, "unknown solve flags");
    *st = msbg_poisson_stats{};
    if (!p->n_active) {
        st->converged = 1;
        return MSBG_OK;
    }
    MSBG_CHECK(!prm->learned || prm->phi, MSBG_EINVAL, "the learned correction needs the phase field grid");
    float *lo = prm->x_lo;
    MSBG_CHECK(!lo || prm->refine >= 1, MSBG_EINVAL, "a compensated iterate (x_lo) needs refine = 1 or 2");
    MSBG_CHECK(!lo || !prm->learned, MSBG_EINVAL, "the learned correction takes a plain FP32 iterate (x_lo = NULL)");
    MSBG_CHECK(!lo || lo != x, MSBG_EINVAL, "x_lo must be a field of its own");
    if (prm->phi)
        MSBG_CHECK(prm->phi->L == 4 && prm->phi->sx == p->dom.sx && prm->phi->sy == p->dom.sy &&
                       prm->phi->sz == p->dom.sz,
                   MSBG_EINVAL, "phase field grid must have the layout's dimensions and 16^3 blocks");
    const int mode = prm->refine;
    if (mode) {
        rc = ensure_inner(p, prm->storage);
        if (rc) return rc;
        if (mode == 2) {
            rc = mg::ensure(p);
            if (rc) return rc;
        }
    } else if (!p->d_r) {
        MSBG_HIP(hipMalloc(&p->d_r, (size_t)p->n_slots * 4096 * 4));
    }
    st->levels = mode == 2 ? mg::n_levels(p) : 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    MSBG_HIP(hipEventCreate(&e0));
    MSBG_HIP(hipEventCreate(&e1));
    MSBG_HIP(hipEventRecord(e0, stream()));
    residual_of(p, stream(), x, lo, b, p->d_r);
    MSBG_HIP(hipGetLastError());
    double r2 = 0.0, b2 = 0.0;
    rc = read_norms(p, &r2, &b2);
    st->checks = 1;
    bool corrected = true;  // no correction before the first period
    double prev_rel = HUGE_VAL;
    int stalls = 0;         // consecutive V-cycles that took off less than 10% of the residual
    while (rc == MSBG_OK) {
        if (prm->learned && !corrected) {
            const double rel = b2 > 0.0 ? std::sqrt(r2 / b2) : std::sqrt(r2);
            if (rel > prm->tol && std::isfinite(rel)) {
                int acc = 0;
                rc = learned::correct(p, prm->learned, prm->phi, x, b, r2, &r2, &b2, &acc);
                if (rc) break;
                st->corrections += 1;
                st->accepted += acc;
                st->checks += 1;
            }
        }
        corrected = true;
        const double rel = b2 > 0.0 ? std::sqrt(r2 / b2) : std::sqrt(r2);
        st->r2 = r2;
        st->b2 = b2;
        st->rel_residual = rel;
        if (rel <= prm->tol) {
            st->converged = 1;
            break;
        }
        // A V-cycle divides the residual by ~8 (by ~1.7 on the hardest bands
        // tested); two cycles in a row that take off less than 10% mean the
        // iterate is at its rounding floor.
        if (mode == 2) {
            stalls = rel > 0.9 * prev_rel ? stalls + 1 : 0;
            if (stalls >= 2) break;
            prev_rel = rel;
        }
        // fine-level red-black sweeps of one check period
        const int per = mode == 2 ? 2 * prm->sweeps_per_check : prm->sweeps_per_check;
        if (!std::isfinite(rel) || st->sweeps + per > prm->max_sweeps) break;
        if ((prm->flags & MSBG_SOLVE_DIRECT) || (mode == 2 && p->n_cells >= kDirectCycleCells)) {
            // a cycle of a band this large is milliseconds of GPU work, which
            // hides its ~60 launches; a graph would cost its instantiation
            // (~4 ms) on every new band
            period(p, stream(), mode, prm->storage, x, lo, b, prm->omega, prm->sweeps_per_check);
            hipError_t le = hipGetLastError();
            if (le != hipSuccess) rc = hip_fail(le, "msbg_poisson_solve");
        } else {
            rc = run_period_graph(p, mode, prm->storage, x, lo, b, prm->omega, prm->sweeps_per_check);
        }
        if (rc) break;
        st->sweeps += per;
        st->checks += 1;
        rc = read_norms(p, &r2, &b2);
        corrected = false;
    }
    hipError_t e = hipEventRecord(e1, stream());
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    MSBG_HIP(e);
    st->ms = ms;
    return MSBG_OK;
}

}  // extern "C"

// Geometric multigrid for the Poisson band (msbg_poisson): one V-cycle on the
// error equation A e = rt of an FP32 iterative-refinement step.
//
// The reference has no solver (docs/roadmap.md:610-702 "HAVEN'T STARTED";
// restriction by the mean of 8 children is its downsample_channel_avg,
// multires/downsample.rs:13-58).  The hierarchy is block-aligned and
// cell-centred:
//   level 0        the band itself (16^3 cells per listed block, storage T)
//   levels 1..3    8^3, 4^3, 2^3 cells per listed block (FP32, structured:
//                  cell (x, y, z) of listed block k at k B^3 + x + B (y + B z))
//   level 4        one cell per listed block (index = position in the list)
//   levels 5..     the block lattice coarsened by 2 per axis, a cell existing
//                  where any child exists
// Levels 4 and coarser are FP32 cell lists whose neighbour, coefficient,
// child and interpolation tables are built on the host once per band; the
// coarsest (<= kCoarseMax cells) is solved in FP64 by conjugate gradients in
// one work-group.
//
// Level l (cell edge h = 2^l fine cells) holds the cells whose first fine
// cell lies inside the domain.  Operators (unscaled):
//   levels 1..3:  (A x)_c = sum over faces of c_f (x_c - x_f), c_f = 1 towards
//                 a cell of the level, 0 towards the domain boundary (Neumann,
//                 as the band mirrors there), w_l = 2 / (1 + 2^-l) towards a
//                 cell of an unlisted block (x_f = 0).  The band's Dirichlet
//                 values sit at the centres of the first outside fine cells,
//                 (h + 1) / 2 from a level-l cell centre; w_l is h over that
//                 distance.
//   levels 4..:   a quarter of the Galerkin product P^T A P of the next finer
//                 level with piecewise-constant P: the coefficient between
//                 two cells is a quarter of the sum of the fine links between
//                 their children, and each cell keeps a quarter of its
//                 children's Dirichlet coefficients, scaled by w_l / w_{l-1}
//                 (for whole cells this is the operator above on level l; for
//                 partly filled cells the band's Dirichlet coupling is
//                 carried down, so no level loses its boundary values).
// Restriction: b_{l+1} = 0.5 * (sum of the child residuals) (4 x their mean:
// h_{l+1}^2 / h_l^2 of the unscaled operators).  Prolongation: trilinear
// (weights 27, 9, 3, 1 / 64), a partner outside the band reading 0 and one
// beyond the domain the clamped cell.  Smoothing: red-black Gauss-Seidel /
// SOR with the caller's omega, colour 0 then 1 before the coarse correction
// and 1 then 0 after it.
#include <algorithm>
#include <cmath>
#include <vector>

#include "poisson.hpp"

namespace msbg {
namespace mg {

constexpr int kCoarseMax = 1024;  // cells of the coarsest level (one FP64 CG work-group)
constexpr int kCgMaxIter = 1000;
constexpr double kCgTol = 1e-6;  // relative; the V-cycle needs far less

struct Level {
    int l = 0;         // cell edge 2^l fine cells
    int B = 0;         // cells per block edge (structured levels 1..3), 0 for cell lists
    int64_t n = 0;     // cells
    float w = 1.0f;    // Dirichlet face coefficient (structured levels)
    float *x = nullptr, *b = nullptr;
    // cell lists (level >= 4)
    int32_t *nb = nullptr;    // [n][6]: -x, +x, -y, +y, -z, +z linked cell, -1 none
    float *cf = nullptr;      // [n][6]: coefficient of the link
    float *diag = nullptr;    // sum of the link and Dirichlet coefficients
    uint8_t *col = nullptr;   // (x + y + z) & 1 of the cell's lattice coordinates
    int32_t *ch = nullptr;    // [n][8]: children in the next finer cell list (level >= 5), -1 none
    int32_t *pn = nullptr;    // [n][8]: interpolation partners in the next coarser level, -1 none
    // coarsest level: connected components, and which have no Dirichlet coupling
    int16_t *comp = nullptr;  // [n]: component of each cell
    uint8_t *sing = nullptr;  // [n_comp]: 1 when the component's operator is singular
    int n_comp = 0, n_sing = 0;
};

struct Hierarchy {
    std::vector<Level> lv;    // lv[k] is level k + 1
    int32_t *nb27 = nullptr;  // [n_active][27]: listed index of block k's neighbour (dx, dy, dz), -1 / -2
};

namespace {

constexpr int kT = 256;

__device__ __forceinline__ int nb27_at(int dx, int dy, int dz) { return (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1); }

// Cell (cx, cy, cz) of a listed block with voxel origin o, at cell edge h:
// inside the domain when its first fine cell is.
__device__ __forceinline__ bool live(const PBlock &B, const Dom &d, int h, int cx, int cy, int cz) {
    return B.ox + cx * h < d.sx && B.oy + cy * h < d.sy && B.oz + cz * h < d.sz;
}

// (b - A x) at cell (cx, cy, cz) of listed block k on a structured level of
// B cells per edge (h = 16 / B); *diag_out receives the diagonal.
template <int B>
__device__ __forceinline__ float blk_residual(const float *__restrict__ x, const float *__restrict__ b,
                                              const int32_t *__restrict__ T, const PBlock &Bk, const Dom &d, i64 k,
                                              int cx, int cy, int cz, float w, float *diag_out) {
    constexpr int B3 = B * B * B, h = 16 / B;
    const i64 i = k * B3 + cx + B * (cy + B * cz);
    const int c[3] = {cx, cy, cz};
    // Every face's neighbour cell first, then the neighbour blocks of the
    // faces that leave this block in one batch of loads (they depend on the
    // cell only), then all six values in one batch: two dependent round trips
    // instead of one or two per face.  A load whose value goes unused reads a
    // valid address (the block's centre entry, the block's own cells).
    int qi[6], tix[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        const int a = f >> 1, dir = (f & 1) ? 1 : -1;
        int q[3] = {cx, cy, cz}, o[3] = {0, 0, 0};
        q[a] += dir;
        if (q[a] < 0) {
            q[a] += B;
            o[a] = -1;
        } else if (q[a] >= B) {
            q[a] -= B;
            o[a] = 1;
        }
        qi[f] = q[0] + B * (q[1] + B * q[2]);
        tix[f] = nb27_at(o[0], o[1], o[2]);  // 13 (the block itself) when the face stays inside
    }
    int32_t tv[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) tv[f] = T[tix[f]];
    i64 nbk[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) nbk[f] = tix[f] == 13 ? k : (i64)tv[f];
    float xv[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) xv[f] = x[(nbk[f] >= 0 ? nbk[f] : k) * B3 + qi[f]];
    const float bi = b[i], xc = x[i];
    const int org[3] = {Bk.ox, Bk.oy, Bk.oz}, dom[3] = {d.sx, d.sy, d.sz};
    float s = 0.0f, dg = 0.0f;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        const int a = f >> 1, dir = (f & 1) ? 1 : -1;
        if (dir > 0 && org[a] + (c[a] + 1) * h >= dom[a]) continue;  // beyond the domain: Neumann
        if (nbk[f] >= 0) {
            s += xv[f];
            dg += 1.0f;
        } else if (nbk[f] == -1) {
            dg += w;
        }
    }
    *diag_out = dg;
    return bi + s - dg * xc;
}

// One half sweep over the cells of one colour: thread j of listed block k
// takes row j / (B / 2) (y + B z) and the cell of that row with
// (x + y + z) & 1 == color.
template <int B>
__global__ __launch_bounds__(kT) void relax_blk_kernel(i64 n_active, const PBlock *__restrict__ blocks, Dom d,
                                                       const int32_t *__restrict__ nb27, float *x,
                                                       const float *__restrict__ b, float w, float omega, int color) {
    constexpr int B3 = B * B * B, H = B3 / 2;
    const i64 i = (i64)blockIdx.x * kT + threadIdx.x;
    if (i >= n_active * H) return;
    const i64 k = i / H;
    const int j = (int)(i % H), row = j / (B / 2), cy = row % B, cz = row / B;
    const int cx = 2 * (j % (B / 2)) + ((cy + cz + color) & 1);  // B even: the block origin keeps the parity
    const PBlock Bk = blocks[k];
    if (!live(Bk, d, 16 / B, cx, cy, cz)) return;
    float dg;
    const float t = blk_residual<B>(x, b, nb27 + k * 27, Bk, d, k, cx, cy, cz, w, &dg);
    const i64 c = k * B3 + cx + B * (cy + B * cz);
    if (dg > 0.0f) x[c] += omega * t / dg;  // dg = 0: a lone cell with Neumann faces only
}

// The first half sweep from x = 0, over every cell: colour 0 takes
// omega b / diag, colour 1 (and every cell outside the domain) 0.
template <int B>
__global__ __launch_bounds__(kT) void init_blk_kernel(i64 n_active, const PBlock *__restrict__ blocks, Dom d,
                                                      const int32_t *__restrict__ nb27, float *x,
                                                      const float *__restrict__ b, float w, float omega) {
    constexpr int B3 = B * B * B, h = 16 / B;
    const i64 i = (i64)blockIdx.x * kT + threadIdx.x;
    if (i >= n_active * B3) return;
    const i64 k = i / B3;
    const int v = (int)(i % B3), c[3] = {v % B, (v / B) % B, v / (B * B)};
    const PBlock Bk = blocks[k];
    float out = 0.0f;
    if (((c[0] + c[1] + c[2]) & 1) == 0 && live(Bk, d, h, c[0], c[1], c[2])) {
        const int32_t *T = nb27 + k * 27;
        const int org[3] = {Bk.ox, Bk.oy, Bk.oz}, dom[3] = {d.sx, d.sy, d.sz};
        float dg = 0.0f;
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            const int a = f >> 1, dir = (f & 1) ? 1 : -1;
            if (dir > 0 && org[a] + (c[a] + 1) * h >= dom[a]) continue;
            const int q = c[a] + dir;
            if (q >= 0 && q < B) {
                dg += 1.0f;
                continue;
            }
            const int32_t nbk = T[nb27_at(a == 0 ? dir : 0, a == 1 ? dir : 0, a == 2 ? dir : 0)];
            dg += nbk >= 0 ? 1.0f : (nbk == -1 ? w : 0.0f);
        }
        if (dg > 0.0f) out = omega * b[i] / dg;
    }
    x[i] = out;
}

// b_c = 0.5 * sum of the residuals of its children (structured level of B
// cells per edge into the level of B / 2; B = 2 gives one cell per block).
// Lane 8 j + o evaluates child o of coarse cell j; the 8 lanes of a coarse
// cell sum by shuffles in a fixed order.
template <int B>
__global__ __launch_bounds__(kT) void restrict_blk_kernel(i64 n_active, const PBlock *__restrict__ blocks, Dom d,
                                                          const int32_t *__restrict__ nb27,
                                                          const float *__restrict__ x, const float *__restrict__ b,
                                                          float w, float *bc) {
    constexpr int C = B / 2, C3 = C * C * C;
    const i64 t = (i64)blockIdx.x * kT + threadIdx.x;
    const i64 i = t >> 3;  // coarse cell
    const int o = (int)(t & 7);
    float r = 0.0f;
    if (i < n_active * C3) {
        const i64 k = i / C3;
        const int v = (int)(i % C3), ux = v % C, uy = (v / C) % C, uz = v / (C * C);
        const PBlock Bk = blocks[k];
        const int fx = 2 * ux + (o & 1), fy = 2 * uy + ((o >> 1) & 1), fz = 2 * uz + (o >> 2);
        float dg;
        if (live(Bk, d, 16 / B, fx, fy, fz)) r = blk_residual<B>(x, b, nb27 + k * 27, Bk, d, k, fx, fy, fz, w, &dg);
    }
    r += __shfl_xor(r, 1);
    r += __shfl_xor(r, 2);
    r += __shfl_xor(r, 4);
    if (o == 0 && i < n_active * C3) bc[i] = 0.5f * r;
}

// Trilinear value at fine cell (cx, cy, cz) of listed block k from the
// coarse level of C cells per block edge (coarse cell edge hc = 16 / C).
template <int C>
__device__ __forceinline__ float blk_interp(const float *__restrict__ xc, const int32_t *__restrict__ T,
                                            const PBlock &Bk, const Dom &d, i64 k, int cx, int cy, int cz) {
    constexpr int C3 = C * C * C, hc = 16 / C;
    const int c[3] = {cx, cy, cz}, org[3] = {Bk.ox, Bk.oy, Bk.oz}, dom[3] = {d.sx, d.sy, d.sz};
    // Loads in three batches (the -x / -y / -z neighbour ids, the 8 partners'
    // block ids, the 8 values), each load at a valid address whether or not
    // its value is used; the sum runs over the partners in the same order.
    int32_t tm[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) tm[a] = T[nb27_at(a == 0 ? -1 : 0, a == 1 ? -1 : 0, a == 2 ? -1 : 0)];
    int p[3], dd[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        p[a] = c[a] >> 1;
        dd[a] = (c[a] & 1) ? 1 : -1;
        // a partner beyond the domain is the cell itself (clamp)
        if (dd[a] > 0 && org[a] + (p[a] + 1) * hc >= dom[a]) dd[a] = 0;
        if (dd[a] < 0 && p[a] == 0 && tm[a] == -2) dd[a] = 0;
    }
    int qi[8], tix[8];
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
        int q[3], o[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int on = (kk >> a) & 1;
            q[a] = p[a] + (on ? dd[a] : 0);
            o[a] = 0;
            if (q[a] < 0) {
                q[a] += C;
                o[a] = -1;
            } else if (q[a] >= C) {
                q[a] -= C;
                o[a] = 1;
            }
        }
        qi[kk] = q[0] + C * (q[1] + C * q[2]);
        tix[kk] = nb27_at(o[0], o[1], o[2]);  // 13: the block itself
    }
    int32_t tv[8];
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) tv[kk] = T[tix[kk]];
    i64 nbk[8];
    float xv[8];
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
        nbk[kk] = tix[kk] == 13 ? k : (i64)tv[kk];
        xv[kk] = xc[(nbk[kk] >= 0 ? nbk[kk] : k) * C3 + qi[kk]];
    }
    float v = 0.0f;
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
        float wt = 1.0f;
#pragma unroll
        for (int a = 0; a < 3; ++a) wt *= ((kk >> a) & 1) ? 0.25f : 0.75f;
        if (nbk[kk] >= 0) v += wt * xv[kk];
    }
    return v;
}

template <int B>
__global__ __launch_bounds__(kT) void prolong_blk_kernel(i64 n_active, const PBlock *__restrict__ blocks, Dom d,
                                                         const int32_t *__restrict__ nb27,
                                                         const float *__restrict__ xc, float *x) {
    constexpr int B3 = B * B * B;
    const i64 i = (i64)blockIdx.x * kT + threadIdx.x;
    if (i >= n_active * B3) return;
    const i64 k = i / B3;
    const int v = (int)(i % B3), cx = v % B, cy = (v / B) % B, cz = v / (B * B);
    const PBlock Bk = blocks[k];
    if (!live(Bk, d, 16 / B, cx, cy, cz)) return;
    x[i] += blk_interp<B / 2>(xc, nb27 + k * 27, Bk, d, k, cx, cy, cz);
}

// Level 0 -> 1: b1 = 0.5 * sum of the 8 children of the FP32 residual r
// (slot layout); children outside the domain count 0.
__global__ __launch_bounds__(kT) void restrict0_kernel(i64 n_active, const PBlock *__restrict__ blocks, Dom d,
                                                       const float *__restrict__ r, float *b1) {
    const i64 i = (i64)blockIdx.x * kT + threadIdx.x;
    if (i >= n_active * 512) return;
    const PBlock B = blocks[i >> 9];
    const int v = (int)(i & 511), ux = v & 7, uy = (v >> 3) & 7, uz = v >> 6;
    const i64 base = ((i64)B.slot << 12);
    // the 8 children as 4 x-pairs, all loaded before the sum (same terms, same order)
    float2 pr[4];
#pragma unroll
    for (int h = 0; h < 4; ++h)
        pr[h] = *reinterpret_cast<const float2 *>(r + base + 2 * ux + 16 * (2 * uy + (h & 1)) + 256 * (2 * uz + (h >> 1)));
    float sum = 0.0f;
#pragma unroll
    for (int o = 0; o < 8; ++o) {
        const int fx = 2 * ux + (o & 1), fy = 2 * uy + ((o >> 1) & 1), fz = 2 * uz + (o >> 2);
        if (live(B, d, 1, fx, fy, fz)) sum += (o & 1) ? pr[o >> 1].y : pr[o >> 1].x;
    }
    b1[i] = 0.5f * sum;
}

// Value of level-1 cell (qx, qy, qz) relative to listed block k (coordinates
// may step one block out on each axis); 0 outside the band.
__device__ __forceinline__ float level1_at(const float *__restrict__ x1, const int32_t *__restrict__ T, i64 k, int qx,
                                           int qy, int qz) {
    const int ox = qx < 0 ? -1 : (qx >= 8 ? 1 : 0), oy = qy < 0 ? -1 : (qy >= 8 ? 1 : 0),
              oz = qz < 0 ? -1 : (qz >= 8 ? 1 : 0);
    const i64 nbk = (ox | oy | oz) ? (i64)T[nb27_at(ox, oy, oz)] : k;
    if (nbk < 0) return 0.0f;
    return x1[nbk * 512 + (qx & 7) + 8 * (qy & 7) + 64 * (qz & 7)];
}

// Level 1 -> 0: e += P x1 on the in-domain cells (e in storage ST).  One
// work-group per listed block stages the block's level-1 cells with a 1-cell
// ring (10^3 values, 0 outside the band) in LDS; thread t then interpolates
// the 16 cells of row (y, z) = (t & 15, t >> 4): their trilinear partners are
// the level-1 x-positions -1 .. 8 on 4 (y, z) rows.  A partner beyond the
// domain is the cell itself (clamp), one outside the band reads 0.
template <int ST>
__global__ __launch_bounds__(256) void prolong0_kernel(const PBlock *__restrict__ blocks, Dom d,
                                                       const int32_t *__restrict__ nb27, const float *__restrict__ x1,
                                                       typename Sto<ST>::T *e) {
    using S = Sto<ST>;
    __shared__ float s_c[1000];
    const i64 k = blockIdx.x;
    const PBlock B = blocks[k];
    const int32_t *T = nb27 + k * 27;
    for (int s = threadIdx.x; s < 1000; s += 256)
        s_c[s] = level1_at(x1, T, k, s % 10 - 1, (s / 10) % 10 - 1, s / 100 - 1);
    __syncthreads();
    const int cy = threadIdx.x & 15, cz = threadIdx.x >> 4;
    if (B.oy + cy >= d.sy || B.oz + cz >= d.sz) return;
    const int py = cy >> 1, pz = cz >> 1;
    int dy = (cy & 1) ? 1 : -1, dz = (cz & 1) ? 1 : -1;
    if ((dy > 0 && B.oy + 2 * (py + 1) >= d.sy) || (dy < 0 && py == 0 && T[nb27_at(0, -1, 0)] == -2)) dy = 0;
    if ((dz > 0 && B.oz + 2 * (pz + 1) >= d.sz) || (dz < 0 && pz == 0 && T[nb27_at(0, 0, -1)] == -2)) dz = 0;
    const bool left_wall = T[nb27_at(-1, 0, 0)] == -2;
    const float *rows[4];
#pragma unroll
    for (int yz = 0; yz < 4; ++yz) {
        const int qy = py + ((yz & 1) ? dy : 0), qz = pz + ((yz & 2) ? dz : 0);
        rows[yz] = s_c + (qz + 1) * 100 + (qy + 1) * 10 + 1;  // rows[yz][qx], qx = -1 .. 8
    }
    float out[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int px = i >> 1;
        int qn = px + ((i & 1) ? 1 : -1);
        if ((qn > px && B.ox + 2 * qn >= d.sx) || (qn < 0 && left_wall)) qn = px;  // clamp
        float val = 0.0f;
#pragma unroll
        for (int yz = 0; yz < 4; ++yz) {
            const float wyz = ((yz & 1) ? 0.25f : 0.75f) * ((yz & 2) ? 0.25f : 0.75f);
            val += wyz * (0.75f * rows[yz][px] + 0.25f * rows[yz][qn]);
        }
        out[i] = val;
    }
    const i64 j = ((i64)B.slot << 12) + 16 * cy + 256 * cz;
    const int n = min(16, d.sx - B.ox);
    if (n == 16) {
        if constexpr (ST == MSBG_STORAGE_F32) {
            float4 *q = reinterpret_cast<float4 *>(e + j);
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                float4 a = q[m];
                a.x += out[4 * m];
                a.y += out[4 * m + 1];
                a.z += out[4 * m + 2];
                a.w += out[4 * m + 3];
                q[m] = a;
            }
        } else {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                uint4 q = *reinterpret_cast<const uint4 *>(e + j + 8 * h);
                uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const float lo = S::ld(reinterpret_cast<const typename S::T *>(&w[m]), 0) + out[8 * h + 2 * m];
                    const float hi = S::ld(reinterpret_cast<const typename S::T *>(&w[m]), 1) + out[8 * h + 2 * m + 1];
                    w[m] = (uint32_t)S::cv(lo) | (uint32_t)S::cv(hi) << 16;
                }
                *reinterpret_cast<uint4 *>(e + j + 8 * h) = make_uint4(w[0], w[1], w[2], w[3]);
            }
        }
    } else {
        for (int i = 0; i < n; ++i) e[j + i] = S::cv(S::ld(e, j + i) + out[i]);
    }
}

__device__ __forceinline__ float list_residual(const float *__restrict__ x, const float *__restrict__ b,
                                               const int32_t *__restrict__ nb, const float *__restrict__ cf,
                                               const float *__restrict__ diag, i64 i) {
    // neighbour indices and coefficients, then the neighbour values, each in
    // one batch (an absent neighbour reads the cell itself, unused)
    int32_t q[6];
    float c[6], xv[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        q[f] = nb[i * 6 + f];
        c[f] = cf[i * 6 + f];
    }
#pragma unroll
    for (int f = 0; f < 6; ++f) xv[f] = x[q[f] >= 0 ? (i64)q[f] : i];
    float s = 0.0f;
#pragma unroll
    for (int f = 0; f < 6; ++f)
        if (q[f] >= 0) s += c[f] * xv[f];
    return b[i] + s - diag[i] * x[i];
}

__global__ __launch_bounds__(kT) void relax_list_kernel(i64 n, const int32_t *__restrict__ nb,
                                                        const float *__restrict__ cf, const float *__restrict__ diag,
                                                        const uint8_t *__restrict__ col, float *x,
                                                        const float *__restrict__ b, float omega, int color) {
    const i64 i = (i64)blockIdx.x * kT + threadIdx.x;
    if (i >= n || col[i] != color || !(diag[i] > 0.0f)) return;
    x[i] += omega * list_residual(x, b, nb, cf, diag, i) / diag[i];
}

__global__ __launch_bounds__(kT) void init_list_kernel(i64 n, const float *__restrict__ diag,
                                                       const uint8_t *__restrict__ col, float *x,
                                                       const float *__restrict__ b, float omega) {
    const i64 i = (i64)blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    x[i] = (col[i] == 0 && diag[i] > 0.0f) ? omega * b[i] / diag[i] : 0.0f;
}

__global__ __launch_bounds__(kT) void restrict_list_kernel(i64 nc, const int32_t *__restrict__ ch,
                                                           const int32_t *__restrict__ nb,
                                                           const float *__restrict__ cf,
                                                           const float *__restrict__ diag,
                                                           const float *__restrict__ x, const float *__restrict__ b,
                                                           float *bc) {
    const i64 i = (i64)blockIdx.x * kT + threadIdx.x;
    if (i >= nc) return;
    int32_t c[8];
#pragma unroll
    for (int o = 0; o < 8; ++o) c[o] = ch[i * 8 + o];
    // every child's residual is evaluated (an absent child at cell 0, unused),
    // so no branch separates one child's loads from the next one's
    float r[8];
#pragma unroll
    for (int o = 0; o < 8; ++o) r[o] = list_residual(x, b, nb, cf, diag, c[o] >= 0 ? c[o] : 0);
    float sum = 0.0f;
#pragma unroll
    for (int o = 0; o < 8; ++o)
        if (c[o] >= 0) sum += r[o];
    bc[i] = 0.5f * sum;
}

__global__ __launch_bounds__(kT) void prolong_list_kernel(i64 n, const int32_t *__restrict__ pn,
                                                          const float *__restrict__ xc, float *x) {
    const i64 i = (i64)blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    int32_t q[8];
    float xv[8];
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) q[kk] = pn[i * 8 + kk];
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) xv[kk] = xc[q[kk] >= 0 ? q[kk] : 0];  // an absent partner reads entry 0, unused
    float v = 0.0f;
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
        const float wt = ((kk & 1) ? 0.25f : 0.75f) * ((kk & 2) ? 0.25f : 0.75f) * ((kk & 4) ? 0.25f : 0.75f);
        if (q[kk] >= 0) v += wt * xv[kk];
    }
    x[i] += v;
}

// Sum over the work-group in a fixed order (lane butterfly, then the waves
// in order); every thread receives the same value.
__device__ __forceinline__ double wg_sum(double v, double *s_red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += s_red[w];
    __syncthreads();
    return t;
}

// Removes, on each singular component (no Dirichlet coupling: the constants
// are its null space), the component's mean from v.  Thread c < n_comp sums
// component c in cell order (deterministic).
__device__ void project_singular(double &v, bool in, int n, int n_comp, const int16_t *s_comp,
                                 const uint8_t *__restrict__ sing, double *s_val, double *s_mean) {
    const int t = threadIdx.x;
    s_val[t] = in ? v : 0.0;
    __syncthreads();
    if (t < n_comp) {
        double sum = 0.0;
        int cnt = 0;
        if (sing[t])
            for (int i = 0; i < n; ++i)
                if (s_comp[i] == t) {
                    sum += s_val[i];
                    ++cnt;
                }
        s_mean[t] = cnt ? sum / cnt : 0.0;
    }
    __syncthreads();
    if (in && sing[s_comp[t]]) v -= s_mean[s_comp[t]];
    __syncthreads();
}

// Coarsest level: conjugate gradients in FP64, one cell per thread (the
// group holds the cells rounded up to whole waves, <= kCoarseMax), until
// ||r|| <= kCgTol ||b|| or kCgMaxIter iterations (the loop condition is the
// same on every thread, so every barrier is reached by the whole group).  On
// a component without Dirichlet coupling the operator is singular: b is
// projected onto its range first and x onto the zero-mean space after, so
// an inconsistent right-hand side (or rounding) cannot drive x along the
// constants.
__global__ __launch_bounds__(kCoarseMax) void coarse_cg_kernel(int n, const int32_t *__restrict__ nb,
                                                               const float *__restrict__ cf,
                                                               const float *__restrict__ diag,
                                                               const float *__restrict__ b, float *x,
                                                               const int16_t *__restrict__ comp,
                                                               const uint8_t *__restrict__ sing, int n_comp,
                                                               int n_sing) {
    __shared__ double s_p[kCoarseMax];
    __shared__ double s_mean[kCoarseMax];
    __shared__ int16_t s_comp[kCoarseMax];
    __shared__ double s_red[kCoarseMax / 64];
    const int t = threadIdx.x;
    const bool in = t < n;
    int32_t q[6] = {-1, -1, -1, -1, -1, -1};
    double c[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double dg = 0.0, r = 0.0;
    if (in) {
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            q[f] = nb[t * 6 + f];
            c[f] = (double)cf[t * 6 + f];
        }
        dg = (double)diag[t];
        r = (double)b[t];
    }
    if (n_sing > 0) {
        s_comp[t] = in ? comp[t] : (int16_t)0;
        project_singular(r, in, n, n_comp, s_comp, sing, s_p, s_mean);
    }
    double xi = 0.0, p = r;
    double rr = wg_sum(r * r, s_red);
    const double stop = kCgTol * kCgTol * rr;
    for (int it = 0; it < kCgMaxIter && rr > stop; ++it) {
        s_p[t] = p;
        __syncthreads();
        double ap = 0.0;
        if (in) {
            ap = dg * p;
#pragma unroll
            for (int f = 0; f < 6; ++f)
                if (q[f] >= 0) ap -= c[f] * s_p[q[f]];
        }
        const double pap = wg_sum(p * ap, s_red);
        if (!(pap > 0.0)) break;
        const double alpha = rr / pap;
        xi += alpha * p;
        r -= alpha * ap;
        const double rr1 = wg_sum(r * r, s_red);
        p = r + (rr1 / rr) * p;
        rr = rr1;
    }
    if (n_sing > 0) project_singular(xi, in, n, n_comp, s_comp, sing, s_p, s_mean);
    if (in) x[t] = (float)xi;
}

int grid_of(i64 n) { return (int)((n + kT - 1) / kT); }

// nu red-black sweeps (colour 0 then 1; rev: 1 then 0); from_zero: x = 0 on
// entry, the first half sweep written without reading x.
template <int B>
void smooth_blk(const msbg_poisson *p, const Level &L, hipStream_t s, float omega, int nu, bool rev, bool from_zero) {
    const int g = grid_of(L.n / 2);
    for (int it = 0; it < nu; ++it)
        for (int h = 0; h < 2; ++h) {
            if (from_zero && it == 0 && h == 0)
                init_blk_kernel<B><<<grid_of(L.n), kT, 0, s>>>(p->n_active, p->d_blocks, p->dom, p->mg->nb27, L.x,
                                                                L.b, L.w, omega);
            else
                relax_blk_kernel<B><<<g, kT, 0, s>>>(p->n_active, p->d_blocks, p->dom, p->mg->nb27, L.x, L.b, L.w,
                                                     omega, rev ? 1 - h : h);
        }
}

void smooth(const msbg_poisson *p, const Level &L, hipStream_t s, float omega, int nu, bool rev, bool from_zero) {
    if (L.B == 8)
        smooth_blk<8>(p, L, s, omega, nu, rev, from_zero);
    else if (L.B == 4)
        smooth_blk<4>(p, L, s, omega, nu, rev, from_zero);
    else if (L.B == 2)
        smooth_blk<2>(p, L, s, omega, nu, rev, from_zero);
    else
        for (int it = 0; it < nu; ++it)
            for (int h = 0; h < 2; ++h) {
                if (from_zero && it == 0 && h == 0)
                    init_list_kernel<<<grid_of(L.n), kT, 0, s>>>(L.n, L.diag, L.col, L.x, L.b, omega);
                else
                    relax_list_kernel<<<grid_of(L.n), kT, 0, s>>>(L.n, L.nb, L.cf, L.diag, L.col, L.x, L.b, omega,
                                                                  rev ? 1 - h : h);
            }
}

void restrict_to(const msbg_poisson *p, const Level &F, const Level &Cl, hipStream_t s) {
    const i64 na = p->n_active;
    const int32_t *T = p->mg->nb27;
    const int g8 = grid_of(8 * Cl.n);  // a lane per child
    if (F.B == 8)
        restrict_blk_kernel<8><<<g8, kT, 0, s>>>(na, p->d_blocks, p->dom, T, F.x, F.b, F.w, Cl.b);
    else if (F.B == 4)
        restrict_blk_kernel<4><<<g8, kT, 0, s>>>(na, p->d_blocks, p->dom, T, F.x, F.b, F.w, Cl.b);
    else if (F.B == 2)
        restrict_blk_kernel<2><<<g8, kT, 0, s>>>(na, p->d_blocks, p->dom, T, F.x, F.b, F.w, Cl.b);
    else
        restrict_list_kernel<<<grid_of(Cl.n), kT, 0, s>>>(Cl.n, Cl.ch, F.nb, F.cf, F.diag, F.x, F.b, Cl.b);
}

void prolong_to(const msbg_poisson *p, const Level &Cl, const Level &F, hipStream_t s) {
    const i64 na = p->n_active;
    const int32_t *T = p->mg->nb27;
    if (F.B == 8)
        prolong_blk_kernel<8><<<grid_of(F.n), kT, 0, s>>>(na, p->d_blocks, p->dom, T, Cl.x, F.x);
    else if (F.B == 4)
        prolong_blk_kernel<4><<<grid_of(F.n), kT, 0, s>>>(na, p->d_blocks, p->dom, T, Cl.x, F.x);
    else if (F.B == 2)
        prolong_blk_kernel<2><<<grid_of(F.n), kT, 0, s>>>(na, p->d_blocks, p->dom, T, Cl.x, F.x);
    else
        prolong_list_kernel<<<grid_of(F.n), kT, 0, s>>>(F.n, F.pn, Cl.x, F.x);
}

// ---- host construction -----------------------------------------------------

// A level as a cell list on its lattice (cell edge 2^l fine cells), with its
// operator: links (neighbour + coefficient) and Dirichlet coefficients.
struct HostOp {
    int l = 0;
    int64_t dims[3] = {0, 0, 0};
    std::vector<int32_t> pos;    // [n][3] lattice coordinates
    std::vector<uint8_t> alive;  // the cell lies inside the domain
    std::vector<int32_t> nb;     // [n][6] linked cell, -1 none
    std::vector<float> cf;       // [n][6]
    std::vector<float> dir;      // [n] sum of the Dirichlet coefficients
    std::vector<int32_t> dense;  // lattice -> live cell, -1
    int64_t at(int64_t x, int64_t y, int64_t z) const { return x + dims[0] * (y + dims[1] * z); }
    size_t n() const { return alive.size(); }
};

float dirichlet_coef(int l) { return (float)(2.0 / (1.0 + std::ldexp(1.0, -l))); }

// Lattice position q (level l) inside the domain: its first fine cell is.
bool in_domain(const msbg_poisson *p, int l, const int64_t q[3]) {
    const int64_t s[3] = {p->dom.sx, p->dom.sy, p->dom.sz};
    for (int a = 0; a < 3; ++a)
        if (q[a] < 0 || (q[a] << l) >= s[a]) return false;
    return true;
}

// Level 3 (2^3 cells per listed block) as a cell list, index k 8 + x + 2 y + 4 z.
void level3_op(const msbg_poisson *p, HostOp &H) {
    const i64 na = p->n_active;
    H.l = 3;
    H.dims[0] = 2 * p->nx;
    H.dims[1] = 2 * p->ny;
    H.dims[2] = 2 * p->nz;
    H.pos.assign((size_t)na * 24, 0);
    H.alive.assign((size_t)na * 8, 0);
    H.dense.assign((size_t)(H.dims[0] * H.dims[1] * H.dims[2]), -1);
    for (i64 k = 0; k < na; ++k) {
        const int64_t bid = p->bids[k];
        const int64_t bx = bid % p->nx, by = (bid / p->nx) % p->ny, bz = bid / (p->nx * p->ny);
        for (int c = 0; c < 8; ++c) {
            const size_t i = (size_t)k * 8 + c;
            const int64_t q[3] = {2 * bx + (c & 1), 2 * by + ((c >> 1) & 1), 2 * bz + (c >> 2)};
            for (int a = 0; a < 3; ++a) H.pos[i * 3 + a] = (int32_t)q[a];
            if (in_domain(p, 3, q)) {
                H.alive[i] = 1;
                H.dense[H.at(q[0], q[1], q[2])] = (int32_t)i;
            }
        }
    }
    const float w = dirichlet_coef(3);
    H.nb.assign(H.n() * 6, -1);
    H.cf.assign(H.n() * 6, 0.0f);
    H.dir.assign(H.n(), 0.0f);
    for (size_t i = 0; i < H.n(); ++i) {
        if (!H.alive[i]) continue;
        for (int f = 0; f < 6; ++f) {
            int64_t q[3] = {H.pos[i * 3], H.pos[i * 3 + 1], H.pos[i * 3 + 2]};
            q[f >> 1] += (f & 1) ? 1 : -1;
            if (!in_domain(p, 3, q)) continue;  // Neumann
            const int32_t j = H.dense[H.at(q[0], q[1], q[2])];
            if (j >= 0) {
                H.nb[i * 6 + f] = j;
                H.cf[i * 6 + f] = 1.0f;
            } else {
                H.dir[i] += w;
            }
        }
    }
}

// The next coarser level of F (Galerkin / 4) and, with `tables`, its child
// table and F's interpolation table (level 4 needs neither: its transfers are
// the structured kernels).  Coarse cells are numbered in the order of their
// first child, so level 4's cell k is listed block k.
void aggregate(const msbg_poisson *p, const HostOp &F, HostOp &Cl, std::vector<int32_t> &ch,
               std::vector<int32_t> &pn, bool tables = true) {
    Cl.l = F.l + 1;
    for (int a = 0; a < 3; ++a) Cl.dims[a] = (F.dims[a] + 1) >> 1;
    Cl.dense.assign((size_t)(Cl.dims[0] * Cl.dims[1] * Cl.dims[2]), -1);
    Cl.pos.clear();
    std::vector<int32_t> parent(F.n(), -1);
    for (size_t i = 0; i < F.n(); ++i) {
        if (!F.alive[i]) continue;
        const int64_t q[3] = {F.pos[i * 3] >> 1, F.pos[i * 3 + 1] >> 1, F.pos[i * 3 + 2] >> 1};
        int32_t &c = Cl.dense[Cl.at(q[0], q[1], q[2])];
        if (c < 0) {
            c = (int32_t)(Cl.pos.size() / 3);
            for (int a = 0; a < 3; ++a) Cl.pos.push_back((int32_t)q[a]);
        }
        parent[i] = c;
    }
    const size_t nc = Cl.pos.size() / 3;
    Cl.alive.assign(nc, 1);
    Cl.nb.assign(nc * 6, -1);
    Cl.cf.assign(nc * 6, 0.0f);
    Cl.dir.assign(nc, 0.0f);
    const float wr = dirichlet_coef(Cl.l) / dirichlet_coef(F.l);
    for (size_t i = 0; i < F.n(); ++i) {
        if (!F.alive[i]) continue;
        const int32_t c = parent[i];
        Cl.dir[c] += 0.25f * wr * F.dir[i];
        for (int f = 0; f < 6; ++f) {
            const int32_t j = F.nb[i * 6 + f];
            if (j < 0 || parent[j] == c) continue;
            Cl.nb[(size_t)c * 6 + f] = parent[j];
            Cl.cf[(size_t)c * 6 + f] += 0.25f * F.cf[i * 6 + f];
        }
    }
    ch.clear();
    pn.clear();
    if (!tables) return;
    ch.assign(nc * 8, -1);
    for (size_t c = 0; c < nc; ++c)
        for (int o = 0; o < 8; ++o) {
            const int64_t x = 2 * (int64_t)Cl.pos[c * 3] + (o & 1), y = 2 * (int64_t)Cl.pos[c * 3 + 1] + ((o >> 1) & 1),
                          z = 2 * (int64_t)Cl.pos[c * 3 + 2] + (o >> 2);
            if (x < F.dims[0] && y < F.dims[1] && z < F.dims[2]) ch[c * 8 + o] = F.dense[F.at(x, y, z)];
        }
    pn.assign(F.n() * 8, -1);
    for (size_t i = 0; i < F.n(); ++i) {
        if (!F.alive[i]) continue;
        for (int kk = 0; kk < 8; ++kk) {
            int64_t q[3];
            for (int a = 0; a < 3; ++a) {
                const int64_t c = F.pos[i * 3 + a], pc = c >> 1, d = (c & 1) ? 1 : -1;
                q[a] = ((kk >> a) & 1) ? pc + d : pc;
            }
            for (int a = 0; a < 3; ++a) {  // a partner beyond the domain is the clamped cell
                int64_t axis[3] = {0, 0, 0};
                axis[a] = q[a];
                if (q[a] >= Cl.dims[a] || !in_domain(p, Cl.l, axis)) q[a] = F.pos[i * 3 + a] >> 1;
            }
            pn[i * 8 + kk] = Cl.dense[Cl.at(q[0], q[1], q[2])];
        }
    }
}

template <class T>
hipError_t upload(T **d, const std::vector<T> &h);

// Connected components of a level's link graph, and which of them have no
// Dirichlet coefficient (singular operator), uploaded into L.
hipError_t components(const HostOp &H, Level &L) {
    const size_t n = H.n();
    std::vector<int16_t> comp(n, -1);
    std::vector<uint8_t> sing;
    std::vector<int32_t> stack;
    for (size_t s0 = 0; s0 < n; ++s0) {
        if (comp[s0] >= 0) continue;
        const int16_t id = (int16_t)sing.size();
        double dir = 0.0;
        stack.assign(1, (int32_t)s0);
        comp[s0] = id;
        while (!stack.empty()) {
            const int32_t i = stack.back();
            stack.pop_back();
            dir += H.dir[i];
            for (int f = 0; f < 6; ++f) {
                const int32_t j = H.nb[(size_t)i * 6 + f];
                if (j >= 0 && comp[j] < 0) {
                    comp[j] = id;
                    stack.push_back(j);
                }
            }
        }
        sing.push_back(dir > 0.0 ? 0 : 1);
    }
    L.n_comp = (int)sing.size();
    L.n_sing = 0;
    for (uint8_t v : sing) L.n_sing += v;
    if (!L.n_sing) return hipSuccess;
    hipError_t e = upload(&L.comp, comp);
    if (e == hipSuccess) e = upload(&L.sing, sing);
    return e;
}

template <class T>
hipError_t upload(T **d, const std::vector<T> &h) {
    if (h.empty()) return hipSuccess;
    hipError_t e = hipMalloc(d, sizeof(T) * h.size());
    if (e == hipSuccess) e = hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice);
    return e;
}

hipError_t upload_list(const HostOp &H, Level &L) {
    const size_t n = H.n();
    std::vector<float> diag(n);
    std::vector<uint8_t> col(n);
    for (size_t i = 0; i < n; ++i) {
        float d = H.dir[i];
        for (int f = 0; f < 6; ++f) d += H.cf[i * 6 + f];
        diag[i] = d;
        col[i] = (uint8_t)((H.pos[i * 3] + H.pos[i * 3 + 1] + H.pos[i * 3 + 2]) & 1);
    }
    hipError_t e = hipMalloc(&L.x, sizeof(float) * std::max<size_t>(1, n));
    if (e == hipSuccess) e = hipMalloc(&L.b, sizeof(float) * std::max<size_t>(1, n));
    if (e == hipSuccess) e = upload(&L.nb, H.nb);
    if (e == hipSuccess) e = upload(&L.cf, H.cf);
    if (e == hipSuccess) e = upload(&L.diag, diag);
    if (e == hipSuccess) e = upload(&L.col, col);
    return e;
}

}  // namespace

void release(msbg_poisson *p) {
    if (!p || !p->mg) return;
    Hierarchy *H = p->mg;
    for (Level &L : H->lv) {
        (void)hipFree(L.x);
        (void)hipFree(L.b);
        (void)hipFree(L.nb);
        (void)hipFree(L.cf);
        (void)hipFree(L.diag);
        (void)hipFree(L.col);
        (void)hipFree(L.ch);
        (void)hipFree(L.pn);
        (void)hipFree(L.comp);
        (void)hipFree(L.sing);
    }
    (void)hipFree(H->nb27);
    delete H;
    p->mg = nullptr;
}

int ensure(msbg_poisson *p) {
    if (p->mg) return MSBG_OK;
    const i64 na = p->n_active;
    Hierarchy *H = new Hierarchy();
    p->mg = H;
    // each listed block's 27 neighbours (listed index, -1 unlisted, -2 beyond the lattice)
    std::vector<int32_t> lidx((size_t)p->n_blocks, -1);
    for (i64 k = 0; k < na; ++k) lidx[p->bids[k]] = (int32_t)k;
    std::vector<int32_t> nb27((size_t)na * 27);
    for (i64 k = 0; k < na; ++k) {
        const int64_t bid = p->bids[k];
        const int64_t bx = bid % p->nx, by = (bid / p->nx) % p->ny, bz = bid / (p->nx * p->ny);
        for (int t = 0; t < 27; ++t) {
            const int64_t x = bx + t % 3 - 1, y = by + (t / 3) % 3 - 1, z = bz + t / 9 - 1;
            int32_t v = -2;
            if (x >= 0 && x < p->nx && y >= 0 && y < p->ny && z >= 0 && z < p->nz) v = lidx[x + p->nx * (y + p->ny * z)];
            nb27[(size_t)k * 27 + t] = v;
        }
    }
    hipError_t e = upload(&H->nb27, nb27);
    for (int l = 1; l <= 3 && e == hipSuccess; ++l) {
        Level L;
        L.l = l;
        L.B = 16 >> l;
        L.n = na * L.B * L.B * L.B;
        L.w = dirichlet_coef(l);
        e = hipMalloc(&L.x, sizeof(float) * L.n);
        if (e == hipSuccess) e = hipMalloc(&L.b, sizeof(float) * L.n);
        H->lv.push_back(L);
    }
    // cell lists: level 4 (one cell per block) and coarser, to <= kCoarseMax cells
    HostOp fine, cur;
    level3_op(p, fine);
    std::vector<int32_t> ch, pn;
    aggregate(p, fine, cur, ch, pn, false);  // level 4; its children are level 3's structured cells
    bool first = true;
    while (e == hipSuccess) {
        Level L;
        L.l = cur.l;
        L.n = (i64)cur.n();
        e = upload_list(cur, L);
        if (e == hipSuccess && !first) e = upload(&L.ch, ch);
        const bool last = L.n <= kCoarseMax;
        if (last && e == hipSuccess) e = components(cur, L);
        HostOp next;
        if (!last && e == hipSuccess) {
            aggregate(p, cur, next, ch, pn);
            e = upload(&L.pn, pn);
        }
        H->lv.push_back(L);
        if (last) break;
        cur = std::move(next);
        first = false;
    }
    if (e != hipSuccess) {
        release(p);
        return hip_fail(e, "multigrid hierarchy");
    }
    return MSBG_OK;
}

int n_levels(const msbg_poisson *p) { return p->mg ? (int)p->mg->lv.size() + 1 : 0; }

void vcycle(msbg_poisson *p, hipStream_t s, int storage, float omega, int nu, bool first_done) {
    const Hierarchy &H = *p->mg;
    const i64 na = p->n_active;
    const int nl = (int)H.lv.size();
    pband::rb_sweeps(p, s, storage, p->d_e, p->d_rt, omega, nu, false, first_done);
    pband::residual_field(p, s, storage, p->d_e, p->d_rt, p->d_r);
    restrict0_kernel<<<grid_of(na * 512), kT, 0, s>>>(na, p->d_blocks, p->dom, p->d_r, H.lv[0].b);
    for (int k = 0; k + 1 < nl; ++k) {
        const Level &L = H.lv[k];
        smooth(p, L, s, omega, nu, false, true);
        restrict_to(p, L, H.lv[k + 1], s);
    }
    const Level &Lc = H.lv[nl - 1];
    const int cg_threads = (int)std::min<i64>(kCoarseMax, std::max<i64>(64, (Lc.n + 63) / 64 * 64));
    coarse_cg_kernel<<<1, cg_threads, 0, s>>>((int)Lc.n, Lc.nb, Lc.cf, Lc.diag, Lc.b, Lc.x, Lc.comp, Lc.sing,
                                              Lc.n_comp, Lc.n_sing);
    for (int k = nl - 2; k >= 0; --k) {
        const Level &L = H.lv[k];
        prolong_to(p, H.lv[k + 1], L, s);
        smooth(p, L, s, omega, nu, true, false);
    }
    const int g = (int)na;  // a work-group per listed block
    if (storage == MSBG_STORAGE_F32)
        prolong0_kernel<MSBG_STORAGE_F32><<<g, 256, 0, s>>>(p->d_blocks, p->dom, H.nb27, H.lv[0].x, (float *)p->d_e);
    else if (storage == MSBG_STORAGE_F16)
        prolong0_kernel<MSBG_STORAGE_F16><<<g, 256, 0, s>>>(p->d_blocks, p->dom, H.nb27, H.lv[0].x, (uint16_t *)p->d_e);
    else
        prolong0_kernel<MSBG_STORAGE_BF16><<<g, 256, 0, s>>>(p->d_blocks, p->dom, H.nb27, H.lv[0].x,
                                                             (uint16_t *)p->d_e);
    pband::rb_sweeps(p, s, storage, p->d_e, p->d_rt, omega, nu, true);
}

}  // namespace mg
}  // namespace msbg

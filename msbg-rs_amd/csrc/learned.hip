// Learned error correction for the Poisson band: a 4-layer 3^3 convolution
// network on 16^3 blocks with a 2-cell halo, BF16 operands on the MFMA
// matrix cores with FP32 accumulation, trained on the GPU by the kernels
// below (forward, backward-data, backward-weights, Adam).  No reference
// counterpart: the reference has no learned component.
//
// Network (per block): X0 = 20^3 x 2 (scaled residual, phase field),
//   L1 valid 20^3 x 2 -> 18^3 x 16, ReLU     L2 valid 18^3 -> 16^3 x 16, ReLU
//   L3 same 16^3 -> 16^3 x 16, ReLU          L4 same 16^3 -> 16^3 x 1
// out[p][o] = bias[o] + sum_t sum_c W[t][c][o] in[p + off + d_t][c], tap
// t = dx + 3 dy + 9 dz, off = 0 (valid) or -1 (same, zero padding).
// Parameters (FP32 master copy): per layer W[27][16][16] + bias[16], padded
// (layer 1 uses c < 2, layer 4 uses o < 1; the padding stays zero because
// its gradient is zero).  Tensors are channels-last [tile][z][y][x][c], BF16.
//
// Every convolution -- forward, and the data gradient (the same kernel with
// the flipped, channel-transposed weights and offset -off-2) -- is an implicit
// GEMM on the 16x16x32 BF16 MFMA: 16 output cells x 16 channels per MFMA
// chain of 14 tap pairs (K = 16 input channels x 2 taps), the input planes
// staged in LDS.  The weight gradient is a second GEMM with K running over
// cells.
#include <cmath>
#include <vector>

#include "poisson.hpp"

namespace msbg {
namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef uint32_t u4v __attribute__((ext_vector_type(4)));

constexpr int kLP = 27 * 16 * 16 + 16;  // parameters per layer (padded)
constexpr int kParams = 4 * kLP;
constexpr int kTap = 27 * 16 * 16;

__device__ __forceinline__ uint16_t to_bf16(float v) {
    uint32_t u = __float_as_uint(v);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// ---------------------------------------------------------------------------
// Convolution as an implicit GEMM on gfx950's 16x16x32 BF16 MFMA.  A
// work-group (8 waves) computes kZS output z-planes of one tile: it stages the
// kZS + 2 input planes the 3^3 taps reach into LDS as SP x SP planes
// (SP = sout + 2, zero where the tap falls outside the input), so every tap
// read is in bounds and no lane tests a coordinate; then each wave takes kCG
// groups of 16 output cells at a time through the 14 tap pairs (K = 32 = 16
// input channels x 2 taps; the 27th tap pairs with zeros) or, for the
// 2-channel input layer, 2 MFMAs of 16 taps x 2 channels.  Lane layout of
// v_mfma_f32_16x16x32_bf16: A[m][k] and B[k][n] are held as 8 consecutive k by
// lane m (n) + 16 (k / 8), D[m][n] as element m % 4 of lane n + 16 (m / 4).  A
// is the weights (M = output channels), B the input cells (N = 16 cells), so
// lane c + 16 j ends with channels 4 j .. 4 j + 3 of cell c and stores them as
// one 8-byte vector (a wave writes 512 contiguous bytes).  The weight operands
// are arranged per lane by prep_wb_kernel and kept in registers for the whole
// work-group.  Each group's result is one MFMA chain in tap-pair order, so its
// bits do not depend on kZS, kCG or the work-group size.
// ---------------------------------------------------------------------------
// Tiles per inference pass: 512 tiles' activations (163 MB) stay in the
// 256 MB Infinity Cache from one layer to the next; 4096-tile passes stream
// them through HBM and measured 30% slower per tile.
constexpr int64_t kForwardChunk = 512;
constexpr int kZS = 4;   // output z-planes per work-group
// Two 16-cell groups per wave pass: with two MFMA chains the compiler keeps
// the next tap pair's B operands in flight (106 VGPRs, 4 waves per SIMD); with
// four it runs each LDS read and its MFMA back to back to stay within 128.
constexpr int kCG = 2;   // 16-cell groups per wave pass
constexpr int kConvThreads = 512;  // 8 waves share one work-group's staged planes
constexpr int kTP = 14;  // tap pairs of a 16-channel layer
constexpr int kWB = kTP * 64;  // A-operand vectors (16 B) per layer

// Tap of pair tp in lane half h (h = 0: k = 0..15, h = 1: k = 16..31) of a
// 16-channel layer; taps are t = dx + 3 dy + 9 dz, 27 is the zero tap.
// Pairs 0..8 are (dx, dy, 0) / (dx, dy, 1), pairs 9..11 (dx, 0, 2) / (dx, 1, 2),
// pair 12 (0, 2, 2) / (1, 2, 2), pair 13 (2, 2, 2) / zero: within each run of
// pairs the h = 1 tap is the h = 0 tap moved by one plane, one row or one
// cell, so a lane's address is a per-lane base plus a constant per pair.
__host__ __device__ constexpr int tap_of(int tp, int h) {
    return tp < 9 ? tp + 9 * h : tp < 12 ? 18 + (tp - 9) + 3 * h : tp == 12 ? 24 + h : (h ? 27 : 26);
}

struct Conv2Args {
    const uint16_t *in;
    const int32_t *idx;  // tile index into `in` per batch position, or null
    int sin, off;
    const u4v *wb;       // [kTP][64] A operands of this layer
    const float *bias;   // [16] or null
    int relu;
    const uint16_t *mask;  // [tile][sout^3][16]: output zeroed where mask <= 0, or null
    void *out;             // bf16 [tile][sout^3][16], or f32 [tile][sout^3] (channel 0)
    int slabs;             // ceil(sout / kZS)
};

// Output edge sout = SP - 2; input cell (off + x, off + y, off + z) of the
// tile's sin^3 input sits at staged (x, y, z - z0).
template <int CIN, int SP, bool OUTF>
__global__ __launch_bounds__(kConvThreads) void conv2_kernel(Conv2Args a) {
    extern __shared__ u4v lds[];
    constexpr int SOUT = SP - 2, S2 = SOUT * SOUT;
    constexpr int NTP = CIN == 16 ? kTP : 2;
    constexpr int VPC = CIN == 16 ? 2 : 1;  // staged words per cell: 2 x 16 B, or one 4-byte pair
    const int tile = blockIdx.x / a.slabs, slab = blockIdx.x - tile * a.slabs;
    const int z0 = slab * kZS, zn = min(kZS, SOUT - z0);
    const int sin = a.sin, off = a.off;
    const int64_t tin = (int64_t)(a.idx ? a.idx[tile] : tile) * sin * sin * sin;
    {
        // stage planes z0 + off .. z0 + off + zn + 1, rows and columns off .. off + SP - 1
        const int nst = (zn + 2) * SP * SP * VPC;
        for (int i = threadIdx.x; i < nst; i += kConvThreads) {
            const unsigned c = (unsigned)i / VPC, v = (unsigned)i - c * VPC;
            const int p = (int)(c / (SP * SP)), r = (int)(c - p * (SP * SP)), y = r / SP, x = r - y * SP;
            const int iz = z0 + off + p, iy = off + y, ix = off + x;
            const bool in = (unsigned)iz < (unsigned)sin && (unsigned)iy < (unsigned)sin && (unsigned)ix < (unsigned)sin;
            const int64_t g = ((int64_t)iz * sin + iy) * sin + ix;
            if constexpr (CIN == 16) {
                u4v w = {0u, 0u, 0u, 0u};
                if (in) w = reinterpret_cast<const u4v *>(a.in)[(tin + g) * 2 + v];
                lds[i] = w;
            } else {
                uint32_t w = 0u;
                if (in) w = reinterpret_cast<const uint32_t *>(a.in)[tin + g];
                reinterpret_cast<uint32_t *>(lds)[i] = w;
            }
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    u4v wv[NTP];
#pragma unroll
    for (int tp = 0; tp < NTP; ++tp) wv[tp] = a.wb[tp * 64 + lane];
    const int m = lane & 15, kb = lane >> 4, h = kb >> 1;
    // CIN 16: per-lane shifts of the three runs of tap pairs (16-byte words);
    // CIN 2: staged-cell offsets of this lane's 8 taps 16 tp + 4 kb + j (past 26: any in-bounds cell)
    const int dA = h * SP * SP * 2, dB = h * SP * 2, dC = h * 2;
    int toff[CIN == 16 ? 1 : 8];
    if constexpr (CIN != 16) {
#pragma unroll
        for (int tp = 0; tp < 2; ++tp)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int t = 16 * tp + 4 * kb + j;
                toff[tp * 4 + j] = t < 27 ? (t / 9) * SP * SP + ((t / 3) % 3) * SP + t % 3 : 0;
            }
    }
    __syncthreads();
    const int ncell = zn * S2, groups = (ncell + 15) / 16;
    float bias[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) bias[j] = a.bias ? a.bias[4 * kb + j] : 0.0f;
    const uint32_t *L32 = reinterpret_cast<const uint32_t *>(lds);
    constexpr int kWaves = kConvThreads / 64;
    for (int g0 = wave * kCG; g0 < groups; g0 += kWaves * kCG) {
        int base[kCG];
#pragma unroll
        for (int g = 0; g < kCG; ++g) {
            const unsigned q = (unsigned)min((g0 + g) * 16 + m, ncell - 1);  // past the last cell: computed, not stored
            const unsigned cz = q / S2, r = q - cz * S2, cy = r / SOUT, cx = r - cy * SOUT;
            base[g] = (int)((cz * SP + cy) * SP + cx);  // staged cell of tap (0, 0, 0)
            if constexpr (CIN == 16) base[g] = base[g] * 2 + (kb & 1);
        }
        f4 acc[kCG];
#pragma unroll
        for (int g = 0; g < kCG; ++g) acc[g] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int tp = 0; tp < NTP; ++tp) {
            u4v av[kCG];
#pragma unroll
            for (int g = 0; g < kCG; ++g) {
                if constexpr (CIN == 16) {
                    const int t0 = tap_of(tp, 0);
                    const int c0 = 2 * ((t0 / 9) * SP * SP + ((t0 / 3) % 3) * SP + t0 % 3);
                    const int sh = tp < 9 ? dA : tp < 12 ? dB : tp == 12 ? dC : 0;
                    av[g] = lds[base[g] + sh + c0];
                } else {
                    u4v v;
#pragma unroll
                    for (int j = 0; j < 4; ++j) v[j] = L32[base[g] + toff[tp * 4 + j]];
                    av[g] = v;
                }
            }
#pragma unroll
            for (int g = 0; g < kCG; ++g)
                acc[g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, wv[tp]),
                                                                 __builtin_bit_cast(bf8, av[g]), acc[g], 0, 0, 0);
        }
#pragma unroll
        for (int g = 0; g < kCG; ++g) {
            const int grp = g0 + g;
            const int cl = grp * 16 + m;  // this lane's cell; channels 4 kb .. 4 kb + 3
            if (grp >= groups || cl >= ncell) continue;
            const int64_t o = ((int64_t)tile * SOUT + z0) * S2 + cl;
            float v[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                v[j] = acc[g][j] + bias[j];
                if (a.relu) v[j] = v[j] > 0.0f ? v[j] : 0.0f;
            }
            if (a.mask) {
                const uint2 mk = *reinterpret_cast<const uint2 *>(a.mask + o * 16 + 4 * kb);
                const uint16_t mh[4] = {(uint16_t)(mk.x & 0xffffu), (uint16_t)(mk.x >> 16), (uint16_t)(mk.y & 0xffffu),
                                        (uint16_t)(mk.y >> 16)};
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (!(from_bf16(mh[j]) > 0.0f)) v[j] = 0.0f;
            }
            if (OUTF) {
                if (kb == 0) ((float *)a.out)[o] = v[0];
            } else {
                const uint2 w = make_uint2((uint32_t)to_bf16(v[0]) | (uint32_t)to_bf16(v[1]) << 16,
                                           (uint32_t)to_bf16(v[2]) | (uint32_t)to_bf16(v[3]) << 16);
                *reinterpret_cast<uint2 *>((uint16_t *)a.out + o * 16 + 4 * kb) = w;
            }
        }
    }
}

// B operands of the four layers for conv2_kernel, forward (wb) and data
// gradient (wbf: layer l's flipped, channel-transposed weights, layers 1..3):
// lane (n, kb) of tap pair tp holds 8 bf16 at k = 8 kb .. 8 kb + 7.  16-channel
// layers: k = 16 (tap - 2 tp) + c, tap = 2 tp + (kb >> 1); the 2-channel layer
// 0: k = 2 (tap - 16 tp) + c, tap = 16 tp + 4 kb + (i >> 1).  Taps past 26 are 0.
__global__ void prep_wb_kernel(const float *__restrict__ params, u4v *wb, u4v *wbf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 4 * kWB) return;
    const int l = i / kWB, r = i % kWB, tp = r / 64, lane = r % 64;
    const int n = lane & 15, kb = lane >> 4;
    const float *W = params + (int64_t)l * kLP;  // W[t][c][o]
    uint16_t f[8], b[8];
    for (int e = 0; e < 8; ++e) {
        int t, c;
        if (l == 0) {
            t = 16 * tp + 4 * kb + (e >> 1);
            c = e & 1;
            if (tp >= 2) t = 27;
        } else {
            t = tap_of(tp, kb >> 1);
            c = (kb & 1) * 8 + e;
        }
        f[e] = t < 27 ? to_bf16(W[(t * 16 + c) * 16 + n]) : (uint16_t)0;
        b[e] = (t < 27 && l > 0) ? to_bf16(W[((26 - t) * 16 + n) * 16 + c]) : (uint16_t)0;
    }
    u4v vf, vb;
    for (int e = 0; e < 4; ++e) {
        vf[e] = (uint32_t)f[2 * e] | (uint32_t)f[2 * e + 1] << 16;
        vb[e] = (uint32_t)b[2 * e] | (uint32_t)b[2 * e + 1] << 16;
    }
    wb[i] = vf;
    wbf[i] = vb;
}

struct WgArgs {
    const uint16_t *X;
    const int32_t *idx;
    int sin, sout, off;
    const uint16_t *G;  // [tile][sout^3][16]
    int chunks;         // waves per tile
    int gpc;            // 16-cell groups per wave
    int n_tiles;
    float *part;        // [wave][kLP]
};

// Weight / bias gradient partials of one layer: wave (tile, chunk) sums over
// its cells dW[t][c][o] += X[p + off + d_t][c] G[p][o] (MFMA 16x16x16 with
// M = c, N = o, K = 16 cells: A[m][k] from lane m + 16 (k / 4), B[k][n] from
// lane n + 16 (k / 4), D[m][n] in element m % 4 of lane n + 16 (m / 4)) and
// db[o] += G[p][o].
template <int CIN>
__global__ __launch_bounds__(256) void wgrad_kernel(WgArgs a) {
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int tile = wave / a.chunks, chunk = wave - tile * a.chunks;
    if (tile >= a.n_tiles) return;
    const int ncell = a.sout * a.sout * a.sout;
    const int groups = (ncell + 15) / 16;
    const int m = lane & 15, kg = lane >> 4;
    const int64_t xbase = (int64_t)(a.idx ? a.idx[tile] : tile) * a.sin * a.sin * a.sin;
    f4 acc[27];
    for (int t = 0; t < 27; ++t) acc[t] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    float bsum = 0.0f;
    const int g0 = chunk * a.gpc, g1 = min(groups, g0 + a.gpc);
    for (int g = g0; g < g1; ++g) {
        s4 bv;
        int cx[4], cy[4], cz[4];
        bool cvalid[4];
        for (int i = 0; i < 4; ++i) {
            const int c = g * 16 + 4 * kg + i;
            cvalid[i] = c < ncell;
            const int cc = cvalid[i] ? c : 0;
            cx[i] = cc % a.sout;
            cy[i] = (cc / a.sout) % a.sout;
            cz[i] = cc / (a.sout * a.sout);
            const uint16_t gv = cvalid[i] ? a.G[((int64_t)tile * ncell + c) * 16 + m] : (uint16_t)0;
            bv[i] = (short)gv;
            bsum += from_bf16(gv);
        }
        for (int t = 0; t < 27; ++t) {
            s4 av = {0, 0, 0, 0};
            if (m < CIN) {
                for (int i = 0; i < 4; ++i) {
                    const int ix = cx[i] + a.off + t % 3, iy = cy[i] + a.off + (t / 3) % 3,
                              iz = cz[i] + a.off + t / 9;
                    if (cvalid[i] && (unsigned)ix < (unsigned)a.sin && (unsigned)iy < (unsigned)a.sin &&
                        (unsigned)iz < (unsigned)a.sin)
                        av[i] = (short)a.X[(xbase + ((int64_t)iz * a.sin + iy) * a.sin + ix) * CIN + m];
                }
            }
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[t], 0, 0, 0);
        }
    }
    float *P = a.part + (int64_t)wave * kLP;
    for (int t = 0; t < 27; ++t)
        for (int j = 0; j < 4; ++j) P[(t * 16 + 4 * kg + j) * 16 + m] = acc[t][j];
    bsum += __shfl_xor(bsum, 16);
    bsum += __shfl_xor(bsum, 32);
    if (kg == 0) P[kTap + m] = bsum;
}

// grad[k] = sum over the n partials in order (deterministic).
__global__ void reduce_partials_kernel(const float *__restrict__ part, int n, float *grad) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kLP) return;
    float s = 0.0f;
    for (int w = 0; w < n; ++w) s += part[(int64_t)w * kLP + k];
    grad[k] = s;
}

// Loss (y - target)^2 / (2 n 4096): G4 = (y - target) / (n 4096) in channel 0,
// per-work-group partial sums of (y - target)^2.
__global__ __launch_bounds__(256) void loss_kernel(const float *__restrict__ y, const float *__restrict__ target,
                                                   const int32_t *idx, int n, float inv, uint16_t *g4,
                                                   double *part) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double d2 = 0.0;
    if (i < (int64_t)n * 4096) {
        const int b = (int)(i >> 12), c = (int)(i & 4095);
        const float d = y[i] - target[(int64_t)(idx ? idx[b] : b) * 4096 + c];
        d2 = (double)d * (double)d;
        uint16_t *g = g4 + i * 16;
        g[0] = to_bf16(d * inv);
        for (int k = 1; k < 16; ++k) g[k] = 0;
    }
    for (int o = 32; o > 0; o >>= 1) d2 += __shfl_xor(d2, o);
    __shared__ double s[4];
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = d2;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = ((s[0] + s[1]) + s[2]) + s[3];
}

__global__ void loss_reduce_kernel(const double *__restrict__ part, int n, double scale, double *loss_out,
                                   float *loss_f) {
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += part[i];
    *loss_out = s * scale;
    if (loss_f) *loss_f = (float)(s * scale);
}

__global__ void adam_kernel(float *p, float *m, float *v, const float *__restrict__ g, int n, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---- band <-> tiles -------------------------------------------------------
__device__ __forceinline__ int mirror(int g, int s) { return g < 0 ? -g - 1 : (g >= s ? 2 * s - 1 - g : g); }

// X0[k] (20^3 x 2, bf16) of listed block k: channel 0 = r * inv_scale (0 in
// blocks that are not listed), channel 1 = phi; halo cells outside the domain
// mirror (Neumann).  inv_scale = scale[1] if scale is given, else inv_host.
__global__ __launch_bounds__(256) void net_input_kernel(const PBlock *__restrict__ blocks, Dom d, int nx, int ny,
                                                        const int32_t *__restrict__ b2s,
                                                        const uint8_t *__restrict__ listed, const float *r,
                                                        const float *scale, float inv_host, const float *phi_slab,
                                                        const int32_t *__restrict__ phi_b2s, uint16_t *x0, int n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)n * 8000) return;
    const int k = (int)(i / 8000), c = (int)(i % 8000);
    const PBlock B = blocks[k];
    const int gx = mirror(B.ox + c % 20 - 2, d.sx), gy = mirror(B.oy + (c / 20) % 20 - 2, d.sy),
              gz = mirror(B.oz + c / 400 - 2, d.sz);
    const int64_t bid = (gx >> 4) + (int64_t)(gy >> 4) * nx + (int64_t)(gz >> 4) * nx * ny;
    const int vid = (gx & 15) | (gy & 15) << 4 | (gz & 15) << 8;
    const float inv = scale ? scale[1] : inv_host;
    const float rv = listed[bid] ? r[((int64_t)b2s[bid] << 12) + vid] * inv : 0.0f;
    const float pv = phi_slab[((int64_t)phi_b2s[bid] << 12) + vid];
    x0[i * 2] = to_bf16(rv);
    x0[i * 2 + 1] = to_bf16(pv);
}

// target[k][vid] = (xs - x) * inv for the in-domain cells of listed block k, else 0.
__global__ __launch_bounds__(256) void error_target_kernel(const PBlock *__restrict__ blocks, Dom d, const float *x,
                                                           const float *xs, float inv, float *target, int n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)n * 4096) return;
    const int k = (int)(i >> 12), v = (int)(i & 4095);
    const PBlock B = blocks[k];
    const bool in = B.ox + (v & 15) < d.sx && B.oy + ((v >> 4) & 15) < d.sy && B.oz + (v >> 8) < d.sz;
    const int64_t c = ((int64_t)B.slot << 12) + v;
    target[i] = in ? (xs[c] - x[c]) * inv : 0.0f;
}

// delta[slot][vid] = y[k][vid] * out_scale * scale[0] on in-domain cells of listed blocks.
__global__ __launch_bounds__(256) void delta_kernel(const PBlock *__restrict__ blocks, Dom d, const float *y,
                                                    float out_scale, const float *scale, float *delta, int n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)n * 4096) return;
    const int k = (int)(i >> 12), v = (int)(i & 4095);
    const PBlock B = blocks[k];
    const bool in = B.ox + (v & 15) < d.sx && B.oy + ((v >> 4) & 15) < d.sy && B.oz + (v >> 8) < d.sz;
    if (in) delta[((int64_t)B.slot << 12) + v] = (y[i] * out_scale) * scale[0];
}

// q = A delta (delta is zero outside the listed cells); partials of r . q and q . q.
__global__ __launch_bounds__(kThreads) void dots_kernel(const PBlock *__restrict__ blocks, Dom d,
                                                        const float *delta, const float *r, double2 *partial) {
    const PBlock B = blocks[xcd_remap(blockIdx.x, gridDim.x) / kSlices];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int gx = B.ox + tx, gy = B.oy + ty;
    double rq = 0.0, qq = 0.0;
    if (gx < d.sx && gy < d.sy) {
        const int z0 = (xcd_remap(blockIdx.x, gridDim.x) % kSlices) * (16 / kSlices);
        for (int tz = z0; tz < z0 + 16 / kSlices; ++tz) {
            const int gz = B.oz + tz;
            if (gz >= d.sz) break;
            const int vid = tx | ty << 4 | tz << 8;
            const int64_t c = ((int64_t)B.slot << 12) + vid;
            const float xl = tx > 0 ? delta[c - 1] : (gx == 0 ? delta[c] : delta[((int64_t)B.nb[0] << 12) + (vid | 15)]);
            const float xr = gx + 1 >= d.sx ? delta[c]
                                            : (tx < 15 ? delta[c + 1] : delta[((int64_t)B.nb[1] << 12) + (vid & ~15)]);
            const float yd = ty > 0 ? delta[c - 16]
                                    : (gy == 0 ? delta[c] : delta[((int64_t)B.nb[2] << 12) + (vid | 0xf0)]);
            const float yu = gy + 1 >= d.sy
                                 ? delta[c]
                                 : (ty < 15 ? delta[c + 16] : delta[((int64_t)B.nb[3] << 12) + (vid & ~0xf0)]);
            const float zb = tz > 0 ? delta[c - 256]
                                    : (gz == 0 ? delta[c] : delta[((int64_t)B.nb[4] << 12) + (vid | 0xf00)]);
            const float zf = gz + 1 >= d.sz
                                 ? delta[c]
                                 : (tz < 15 ? delta[c + 256] : delta[((int64_t)B.nb[5] << 12) + (vid & ~0xf00)]);
            const float s = ((((xl + xr) + yd) + yu) + zb) + zf;
            const float q = __builtin_fmaf(6.0f, delta[c], -s);
            rq += (double)r[c] * (double)q;
            qq += (double)q * (double)q;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        rq += __shfl_xor(rq, o);
        qq += __shfl_xor(qq, o);
    }
    __shared__ double s_a[kThreads / 64], s_b[kThreads / 64];
    if ((threadIdx.x & 63) == 0) {
        s_a[threadIdx.x >> 6] = rq;
        s_b[threadIdx.x >> 6] = qq;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) {
            a += s_a[w];
            b += s_b[w];
        }
        partial[blockIdx.x] = make_double2(a, b);
    }
}

__global__ __launch_bounds__(256) void reduce_dots_kernel(const double2 *__restrict__ partial, int n, double *out) {
    __shared__ double s_a[256], s_b[256];
    double a = 0.0, b = 0.0;
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
        out[0] = s_a[0];
        out[1] = s_b[0];
    }
}

// x_try = fmaf(alpha, delta, x) on the listed in-domain cells, alpha = (r.q)/(q.q)
// (0 when q = 0); commit = 1 copies x_try back into x instead.
__global__ __launch_bounds__(kThreads) void try_kernel(const PBlock *__restrict__ blocks, Dom d, const double *dots,
                                                       const float *delta, float *x, float *xt, int commit) {
    const PBlock B = blocks[xcd_remap(blockIdx.x, gridDim.x) / kSlices];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int gx = B.ox + tx, gy = B.oy + ty;
    if (gx >= d.sx || gy >= d.sy) return;
    const float alpha = dots[1] > 0.0 ? (float)(dots[0] / dots[1]) : 0.0f;
    const int z0 = (xcd_remap(blockIdx.x, gridDim.x) % kSlices) * (16 / kSlices);
    for (int tz = z0; tz < z0 + 16 / kSlices; ++tz) {
        if (B.oz + tz >= d.sz) break;
        const int64_t c = ((int64_t)B.slot << 12) + (tx | ty << 4 | tz << 8);
        if (commit)
            x[c] = xt[c];
        else
            xt[c] = __builtin_fmaf(alpha, delta[c], x[c]);
    }
}

__global__ void fill_zero_kernel(float *p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

int blocks_for(int64_t threads) { return (int)((threads + 255) / 256); }

}  // namespace
}  // namespace msbg

using namespace msbg;

struct msbg_learned {
    float *d_params = nullptr, *d_m = nullptr, *d_v = nullptr, *d_grad = nullptr;
    msbg::u4v *d_wb = nullptr, *d_wbf = nullptr;  // MFMA B operands, forward and data gradient
    float out_scale = 1.0f;
    int adam_t = 0;
    // workspace for `cap` tiles
    int64_t cap = 0;
    uint16_t *A1 = nullptr, *A2 = nullptr, *A3 = nullptr, *G1 = nullptr, *G2 = nullptr, *G3 = nullptr, *G4 = nullptr;
    uint16_t *X0 = nullptr;  // network input assembled from a Poisson band
    float *Y = nullptr;
    float *wpart = nullptr;
    double *loss_part = nullptr, *d_loss = nullptr;
    int64_t cap_train = 0;
};

namespace msbg {
namespace {

void free_ws(msbg_learned *n) {
    for (uint16_t **p : {&n->A1, &n->A2, &n->A3, &n->G1, &n->G2, &n->G3, &n->G4, &n->X0}) {
        (void)hipFree(*p);
        *p = nullptr;
    }
    (void)hipFree(n->Y);
    (void)hipFree(n->wpart);
    (void)hipFree(n->loss_part);
    n->Y = nullptr;
    n->wpart = nullptr;
    n->loss_part = nullptr;
    n->cap = n->cap_train = 0;
}

// Workspace for `tiles` tiles; `train` adds the backward tensors.
int ensure_ws(msbg_learned *n, int64_t tiles, bool train) {
    if (tiles <= n->cap && (!train || tiles <= n->cap_train)) return MSBG_OK;
    free_ws(n);
    const size_t a18 = (size_t)tiles * 5832 * 16 * 2, a16 = (size_t)tiles * 4096 * 16 * 2;
    MSBG_HIP(hipMalloc(&n->A1, a18));
    MSBG_HIP(hipMalloc(&n->A2, a16));
    MSBG_HIP(hipMalloc(&n->A3, a16));
    MSBG_HIP(hipMalloc(&n->Y, (size_t)tiles * 4096 * 4));
    MSBG_HIP(hipMalloc(&n->X0, (size_t)tiles * 8000 * 2 * 2));
    n->cap = tiles;
    if (train) {
        MSBG_HIP(hipMalloc(&n->G1, a18));
        MSBG_HIP(hipMalloc(&n->G2, a16));
        MSBG_HIP(hipMalloc(&n->G3, a16));
        MSBG_HIP(hipMalloc(&n->G4, a16));
        MSBG_HIP(hipMalloc(&n->wpart, (size_t)tiles * 23 * kLP * 4));
        MSBG_HIP(hipMalloc(&n->loss_part, (size_t)tiles * 16 * sizeof(double)));
        n->cap_train = tiles;
    }
    return MSBG_OK;
}

// Output edge SOUT; the layer reads input cells off .. off + SOUT + 1 of each
// sin^3 tile along every axis (zero outside the tile).
template <int CIN, int SOUT, bool OUTF>
void conv(hipStream_t s, const uint16_t *in, const int32_t *idx, int sin, int off, const u4v *wb,
          const float *bias, int relu, const uint16_t *mask, void *out, int n) {
    Conv2Args a;
    a.in = in;
    a.idx = idx;
    a.sin = sin;
    a.off = off;
    a.wb = wb;
    a.bias = bias;
    a.relu = relu;
    a.mask = mask;
    a.out = out;
    a.slabs = (SOUT + kZS - 1) / kZS;
    constexpr int SP = SOUT + 2;
    const size_t lds = (size_t)(kZS + 2) * SP * SP * CIN * 2;
    conv2_kernel<CIN, SP, OUTF><<<n * a.slabs, kConvThreads, lds, s>>>(a);
}

void forward(msbg_learned *net, hipStream_t s, const uint16_t *x0, const int32_t *idx, int n) {
    const u4v *W = net->d_wb;
    const float *P = net->d_params;
    conv<2, 18, false>(s, x0, idx, 20, 0, W, P + 0 * kLP + kTap, 1, nullptr, net->A1, n);
    conv<16, 16, false>(s, net->A1, nullptr, 18, 0, W + kWB, P + 1 * kLP + kTap, 1, nullptr, net->A2, n);
    conv<16, 16, false>(s, net->A2, nullptr, 16, -1, W + 2 * kWB, P + 2 * kLP + kTap, 1, nullptr, net->A3, n);
    conv<16, 16, true>(s, net->A3, nullptr, 16, -1, W + 3 * kWB, P + 3 * kLP + kTap, 0, nullptr, net->Y, n);
}

template <int CIN>
void wgrad(msbg_learned *net, hipStream_t s, const uint16_t *X, const int32_t *idx, int sin, int sout, int off,
           const uint16_t *G, int n, float *grad) {
    WgArgs a;
    a.X = X;
    a.idx = idx;
    a.sin = sin;
    a.sout = sout;
    a.off = off;
    a.G = G;
    a.gpc = 16;
    a.chunks = ((sout * sout * sout + 15) / 16 + a.gpc - 1) / a.gpc;
    a.n_tiles = n;
    a.part = net->wpart;
    const int64_t waves = (int64_t)a.chunks * n;
    wgrad_kernel<CIN><<<(int)((waves + 3) / 4), 256, 0, s>>>(a);
    reduce_partials_kernel<<<(kLP + 255) / 256, 256, 0, s>>>(net->wpart, (int)waves, grad);
}

// Loss and gradients of n tiles into net->d_grad; loss into net->d_loss (and loss_f if given).
void loss_grad(msbg_learned *net, hipStream_t s, const uint16_t *x0, const float *target, const int32_t *idx, int n,
               float *loss_f) {
    forward(net, s, x0, idx, n);
    const double inv = 1.0 / ((double)n * 4096.0);
    const int nb = n * 16;
    loss_kernel<<<nb, 256, 0, s>>>(net->Y, target, idx, n, (float)inv, net->G4, net->loss_part);
    loss_reduce_kernel<<<1, 64, 0, s>>>(net->loss_part, nb, 0.5 * inv, net->d_loss, loss_f);
    const u4v *Wf = net->d_wbf;
    float *g = net->d_grad;
    wgrad<16>(net, s, net->A3, nullptr, 16, 16, -1, net->G4, n, g + 3 * kLP);
    conv<16, 16, false>(s, net->G4, nullptr, 16, -1, Wf + 3 * kWB, nullptr, 0, net->A3, net->G3, n);
    wgrad<16>(net, s, net->A2, nullptr, 16, 16, -1, net->G3, n, g + 2 * kLP);
    conv<16, 16, false>(s, net->G3, nullptr, 16, -1, Wf + 2 * kWB, nullptr, 0, net->A2, net->G2, n);
    wgrad<16>(net, s, net->A1, nullptr, 18, 16, 0, net->G2, n, g + 1 * kLP);
    conv<16, 18, false>(s, net->G2, nullptr, 16, -2, Wf + 1 * kWB, nullptr, 0, net->A1, net->G1, n);
    wgrad<2>(net, s, x0, idx, 20, 18, 0, net->G1, n, g + 0 * kLP);
}

void prep(msbg_learned *net, hipStream_t s) {
    prep_wb_kernel<<<blocks_for(4 * kWB), 256, 0, s>>>(net->d_params, net->d_wb, net->d_wbf);
}

}  // namespace

namespace learned {

int correct(msbg_poisson *p, msbg_learned *net, const msbg_grid *phi, float *x, const float *b, double r2_in,
            double *r2, double *b2, int *accepted) {
    hipStream_t s = stream();
    const size_t slab = (size_t)p->n_slots * 4096;
    if (!p->d_delta) {
        // all three buffers or none: a failed allocation leaves the band unchanged
        float *delta = nullptr, *xtry = nullptr;
        double *dots = nullptr;
        hipError_t e = hipMalloc(&delta, slab * 4);
        if (e == hipSuccess) e = hipMalloc(&xtry, slab * 4);
        if (e == hipSuccess) e = hipMalloc(&dots, 2 * sizeof(double));
        if (e != hipSuccess) {
            if (delta) (void)hipFree(delta);
            if (xtry) (void)hipFree(xtry);
            return hip_fail(e, "learned correction buffers");
        }
        fill_zero_kernel<<<1024, 256, 0, s>>>(delta, (int64_t)slab);
        p->d_delta = delta;
        p->d_xtry = xtry;
        p->d_dots = dots;
    }
    const int64_t chunk = std::min<int64_t>(p->n_active, kForwardChunk);
    int rc = ensure_ws(net, chunk, false);
    if (rc) return rc;
    uint16_t *x0 = net->X0;
    for (int64_t k0 = 0; k0 < p->n_active; k0 += chunk) {
        const int n = (int)std::min<int64_t>(chunk, p->n_active - k0);
        net_input_kernel<<<blocks_for((int64_t)n * 8000), 256, 0, s>>>(
            p->d_blocks + k0, p->dom, (int)p->nx, (int)p->ny, p->d_b2s, p->d_listed, p->d_r, p->d_scale, 0.0f,
            phi->d_slab, phi->d_b2s, x0, n);
        forward(net, s, x0, nullptr, n);
        delta_kernel<<<blocks_for((int64_t)n * 4096), 256, 0, s>>>(p->d_blocks + k0, p->dom, net->Y, net->out_scale,
                                                                   p->d_scale, p->d_delta, n);
    }
    dots_kernel<<<p->n_wg, kThreads, 0, s>>>(p->d_blocks, p->dom, p->d_delta, p->d_r, p->d_partial);
    reduce_dots_kernel<<<1, 256, 0, s>>>(p->d_partial, p->n_wg, p->d_dots);
    MSBG_HIP(hipMemcpyAsync(p->d_xtry, x, slab * 4, hipMemcpyDeviceToDevice, s));
    try_kernel<<<p->n_wg, kThreads, 0, s>>>(p->d_blocks, p->dom, p->d_dots, p->d_delta, x, p->d_xtry, 0);
    pband::residual(p, s, MSBG_STORAGE_F32, p->d_xtry, b, p->d_r);
    MSBG_HIP(hipGetLastError());
    rc = pband::read_norms(p, r2, b2);
    if (rc) return rc;
    *accepted = *r2 < r2_in;
    if (*accepted) {
        try_kernel<<<p->n_wg, kThreads, 0, s>>>(p->d_blocks, p->dom, p->d_dots, p->d_delta, x, p->d_xtry, 1);
    } else {
        pband::residual(p, s, MSBG_STORAGE_F32, x, b, p->d_r);
        MSBG_HIP(hipGetLastError());
        rc = pband::read_norms(p, r2, b2);
        if (rc) return rc;
    }
    MSBG_HIP(hipGetLastError());
    return MSBG_OK;
}

}  // namespace learned
}  // namespace msbg

extern "C" {

int msbg_learned_param_count(int *n) {
    MSBG_CHECK(n, MSBG_EINVAL, "null argument");
    *n = kParams;
    return MSBG_OK;
}

int msbg_learned_new(msbg_learned **out) {
    MSBG_CHECK(out, MSBG_EINVAL, "null argument");
    msbg_learned *n = new msbg_learned();
    hipError_t e = hipMalloc(&n->d_params, kParams * 4);
    if (e == hipSuccess) e = hipMalloc(&n->d_m, kParams * 4);
    if (e == hipSuccess) e = hipMalloc(&n->d_v, kParams * 4);
    if (e == hipSuccess) e = hipMalloc(&n->d_grad, kParams * 4);
    if (e == hipSuccess) e = hipMalloc(&n->d_wb, 4 * kWB * sizeof(u4v));
    if (e == hipSuccess) e = hipMalloc(&n->d_wbf, 4 * kWB * sizeof(u4v));
    if (e == hipSuccess) e = hipMalloc(&n->d_loss, sizeof(double));
    if (e == hipSuccess) e = hipMemset(n->d_params, 0, kParams * 4);
    if (e == hipSuccess) e = hipMemset(n->d_m, 0, kParams * 4);
    if (e == hipSuccess) e = hipMemset(n->d_v, 0, kParams * 4);
    // the data gradient of layer 1 stages 6 planes of 20^2 x 16 channels (76.8 KB)
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)conv2_kernel<16, 20, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (kZS + 2) * 20 * 20 * 16 * 2);
    if (e != hipSuccess) {
        msbg_learned_free(n);
        return hip_fail(e, "msbg_learned_new");
    }
    prep(n, stream());
    *out = n;
    return MSBG_OK;
}

int msbg_learned_free(msbg_learned *n) {
    if (!n) return MSBG_OK;
    free_ws(n);
    (void)hipFree(n->d_params);
    (void)hipFree(n->d_m);
    (void)hipFree(n->d_v);
    (void)hipFree(n->d_grad);
    (void)hipFree(n->d_wb);
    (void)hipFree(n->d_wbf);
    (void)hipFree(n->d_loss);
    delete n;
    return MSBG_OK;
}

int msbg_learned_set_params(msbg_learned *n, const float *host_params, float out_scale) {
    MSBG_CHECK(n && host_params, MSBG_EINVAL, "null argument");
    MSBG_CHECK(std::isfinite(out_scale), MSBG_EINVAL, "out_scale must be finite");
    MSBG_HIP(hipMemcpyAsync(n->d_params, host_params, kParams * 4, hipMemcpyHostToDevice, stream()));
    MSBG_HIP(hipMemsetAsync(n->d_m, 0, kParams * 4, stream()));
    MSBG_HIP(hipMemsetAsync(n->d_v, 0, kParams * 4, stream()));
    n->adam_t = 0;
    n->out_scale = out_scale;
    prep(n, stream());
    MSBG_HIP(hipGetLastError());
    MSBG_HIP(hipStreamSynchronize(stream()));
    return MSBG_OK;
}

int msbg_learned_get_params(const msbg_learned *n, float *host_params, float *out_scale) {
    MSBG_CHECK(n && host_params && out_scale, MSBG_EINVAL, "null argument");
    MSBG_HIP(hipMemcpyAsync(host_params, n->d_params, kParams * 4, hipMemcpyDeviceToHost, stream()));
    MSBG_HIP(hipStreamSynchronize(stream()));
    *out_scale = n->out_scale;
    return MSBG_OK;
}

int msbg_learned_forward(msbg_learned *n, const uint16_t *d_x0, int64_t count, float *d_y) {
    MSBG_CHECK(n && d_x0 && d_y && count >= 0, MSBG_EINVAL, "bad argument");
    const int64_t chunk = std::min<int64_t>(std::max<int64_t>(count, 1), kForwardChunk);
    int rc = ensure_ws(n, chunk, false);
    if (rc) return rc;
    for (int64_t k0 = 0; k0 < count; k0 += chunk) {
        const int c = (int)std::min<int64_t>(chunk, count - k0);
        forward(n, stream(), d_x0 + k0 * 16000, nullptr, c);
        MSBG_HIP(hipMemcpyAsync(d_y + k0 * 4096, n->Y, (size_t)c * 4096 * 4, hipMemcpyDeviceToDevice, stream()));
    }
    MSBG_HIP(hipGetLastError());
    return MSBG_OK;
}

int msbg_learned_loss_grad(msbg_learned *n, const uint16_t *d_x0, const float *d_target, int64_t count,
                           double *loss, float *host_grad) {
    MSBG_CHECK(n && d_x0 && d_target && loss && host_grad, MSBG_EINVAL, "null argument");
    MSBG_CHECK(count >= 1 && count <= 4096, MSBG_EINVAL, "1 <= count <= 4096 tiles");
    int rc = ensure_ws(n, count, true);
    if (rc) return rc;
    loss_grad(n, stream(), d_x0, d_target, nullptr, (int)count, nullptr);
    MSBG_HIP(hipGetLastError());
    MSBG_HIP(hipMemcpyAsync(host_grad, n->d_grad, kParams * 4, hipMemcpyDeviceToHost, stream()));
    MSBG_HIP(hipMemcpyAsync(loss, n->d_loss, sizeof(double), hipMemcpyDeviceToHost, stream()));
    MSBG_HIP(hipStreamSynchronize(stream()));
    return MSBG_OK;
}

int msbg_learned_train(msbg_learned *n, const uint16_t *d_x0, const float *d_target, int64_t n_pairs, int steps,
                       int batch, float lr, uint64_t seed, float *host_losses) {
    MSBG_CHECK(n && d_x0 && d_target && host_losses, MSBG_EINVAL, "null argument");
    MSBG_CHECK(n_pairs >= 1 && n_pairs < ((int64_t)1 << 31), MSBG_EINVAL, "bad pair count");
    MSBG_CHECK(steps >= 0 && batch >= 1 && batch <= 1024, MSBG_EINVAL, "bad steps / batch");
    if (!steps) return MSBG_OK;
    int rc = ensure_ws(n, batch, true);
    if (rc) return rc;
    hipStream_t s = stream();
    // batch indices: SplitMix64 stream from `seed`, uniform over the pairs
    std::vector<int32_t> idx((size_t)steps * batch);
    uint64_t z = seed;
    for (int32_t &v : idx) {
        z += 0x9e3779b97f4a7c15ULL;
        uint64_t q = z;
        q = (q ^ (q >> 30)) * 0xbf58476d1ce4e5b9ULL;
        q = (q ^ (q >> 27)) * 0x94d049bb133111ebULL;
        q ^= q >> 31;
        v = (int32_t)(q % (uint64_t)n_pairs);
    }
    int32_t *d_idx = nullptr;
    float *d_losses = nullptr;
    MSBG_HIP(hipMalloc(&d_idx, idx.size() * 4));
    hipError_t e = hipMalloc(&d_losses, (size_t)steps * 4);
    if (e == hipSuccess) e = hipMemcpyAsync(d_idx, idx.data(), idx.size() * 4, hipMemcpyHostToDevice, s);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int st = 0; st < steps && e == hipSuccess; ++st) {
        loss_grad(n, s, d_x0, d_target, d_idx + (size_t)st * batch, batch, d_losses + st);
        n->adam_t += 1;
        const float c1 = 1.0f - std::pow(b1, (float)n->adam_t), c2 = 1.0f - std::pow(b2, (float)n->adam_t);
        adam_kernel<<<blocks_for(kParams), 256, 0, s>>>(n->d_params, n->d_m, n->d_v, n->d_grad, kParams, lr, b1, b2,
                                                        eps, c1, c2);
        prep(n, s);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(host_losses, d_losses, (size_t)steps * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d_idx);
    (void)hipFree(d_losses);
    MSBG_HIP(e);
    return MSBG_OK;
}

int msbg_poisson_net_input(msbg_poisson *p, const msbg_grid *phi, const float *d_r, float inv_scale,
                           uint16_t *d_x0) {
    MSBG_CHECK(p && phi && d_r && d_x0, MSBG_EINVAL, "null argument");
    MSBG_CHECK(phi->L == 4 && phi->sx == p->dom.sx && phi->sy == p->dom.sy && phi->sz == p->dom.sz, MSBG_EINVAL,
               "phase field grid must have the layout's dimensions and 16^3 blocks");
    MSBG_CHECK(p->dom.sx >= 2 && p->dom.sy >= 2 && p->dom.sz >= 2, MSBG_EINVAL, "domain must be >= 2 cells per axis");
    if (!p->n_active) return MSBG_OK;
    net_input_kernel<<<blocks_for(p->n_active * 8000), 256, 0, stream()>>>(
        p->d_blocks, p->dom, (int)p->nx, (int)p->ny, p->d_b2s, p->d_listed, d_r, nullptr, inv_scale, phi->d_slab,
        phi->d_b2s, d_x0, (int)p->n_active);
    MSBG_HIP(hipGetLastError());
    return MSBG_OK;
}

int msbg_poisson_error_target(msbg_poisson *p, const float *d_x, const float *d_xstar, float inv_scale,
                              float *d_target) {
    MSBG_CHECK(p && d_x && d_xstar && d_target, MSBG_EINVAL, "null argument");
    if (!p->n_active) return MSBG_OK;
    error_target_kernel<<<blocks_for(p->n_active * 4096), 256, 0, stream()>>>(p->d_blocks, p->dom, d_x, d_xstar,
                                                                             inv_scale, d_target, (int)p->n_active);
    MSBG_HIP(hipGetLastError());
    return MSBG_OK;
}

int msbg_poisson_learned_correct(msbg_poisson *p, msbg_learned *net, const msbg_grid *phi, float *x, const float *b,
                                 double *r2_before, double *r2_after, int *accepted) {
    MSBG_CHECK(p && net && phi && x && b && r2_before && r2_after && accepted, MSBG_EINVAL, "null argument");
    MSBG_CHECK(phi->L == 4 && phi->sx == p->dom.sx && phi->sy == p->dom.sy && phi->sz == p->dom.sz, MSBG_EINVAL,
               "phase field grid must have the layout's dimensions and 16^3 blocks");
    MSBG_CHECK(p->dom.sx >= 2 && p->dom.sy >= 2 && p->dom.sz >= 2, MSBG_EINVAL, "domain must be >= 2 cells per axis");
    *accepted = 0;
    if (!p->n_active) return MSBG_OK;
    if (!p->d_r) MSBG_HIP(hipMalloc(&p->d_r, (size_t)p->n_slots * 4096 * 4));
    pband::residual(p, stream(), MSBG_STORAGE_F32, x, b, p->d_r);
    MSBG_HIP(hipGetLastError());
    double b2 = 0.0;
    int rc = pband::read_norms(p, r2_before, &b2);
    if (rc) return rc;
    return learned::correct(p, net, phi, x, b, *r2_before, r2_after, &b2, accepted);
}

}  // extern "C"

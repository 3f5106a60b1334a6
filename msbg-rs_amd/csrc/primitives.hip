// Device-wide exclusive scan and stable LSD radix sort (wave64 ballot ranking).
#include <algorithm>
#include <vector>

#include "primitives.hpp"

namespace msbg {

// ---------------------------------------------------------------------------
// Workspace.
// ---------------------------------------------------------------------------
static void *g_ws[24] = {nullptr};
static size_t g_ws_size[24] = {0};

void *workspace(int id, size_t bytes, int *rc) {
    *rc = MSBG_OK;
    if (bytes <= g_ws_size[id]) return g_ws[id];
    if (g_ws[id]) {
        (void)hipStreamSynchronize(stream());
        (void)hipFree(g_ws[id]);
        g_ws[id] = nullptr;
        g_ws_size[id] = 0;
    }
    size_t cap = std::max<size_t>(bytes, 1 << 20);
    hipError_t e = hipMalloc(&g_ws[id], cap);
    if (e != hipSuccess) {
        *rc = e == hipErrorOutOfMemory ? fail(MSBG_ENOMEM, "workspace: out of device memory")
                                       : hip_fail(e, "workspace");
        g_ws[id] = nullptr;
        return nullptr;
    }
    g_ws_size[id] = cap;
    return g_ws[id];
}

// ---------------------------------------------------------------------------
// Scan: tiles of 256 lanes x 8 items; tile sums, a one-workgroup scan of the
// sums, then the in-tile scan plus the tile's base.
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 256, kScanItems = 8, kScanTile = kScanThreads * kScanItems;

// Inclusive scan of v across the workgroup (256 lanes, 4 waves); returns the
// inclusive prefix and writes the workgroup total to *total.
__device__ __forceinline__ uint32_t block_scan_inclusive(uint32_t v, uint32_t *s_wave, uint32_t *total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    if (lane == 63) s_wave[wave] = v;
    __syncthreads();
    uint32_t add = 0;
    for (int w = 0; w < wave; ++w) add += s_wave[w];
    if (total) *total = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
    __syncthreads();
    return v + add;
}

__global__ __launch_bounds__(kScanThreads) void scan_tile_sums(const uint32_t *in, int64_t n, uint32_t *sums) {
    __shared__ uint32_t s_wave[4];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) s += in[base + k];
    uint32_t total;
    block_scan_inclusive(s, s_wave, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// One workgroup, exclusive scan of `m` tile sums in place.
__global__ __launch_bounds__(kScanThreads) void scan_sums(uint32_t *sums, int64_t m, uint32_t *grand_total) {
    __shared__ uint32_t s_wave[4];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t off = 0; off < m; off += kScanThreads) {
        const int64_t i = off + threadIdx.x;
        const uint32_t v = i < m ? sums[i] : 0u;
        uint32_t total;
        const uint32_t inc = block_scan_inclusive(v, s_wave, &total);
        const uint32_t c = carry;
        if (i < m) sums[i] = c + inc - v;
        __syncthreads();
        if (threadIdx.x == 0) carry = c + total;
        __syncthreads();
    }
    if (threadIdx.x == 0 && grand_total) *grand_total = carry;
}

__global__ __launch_bounds__(kScanThreads) void scan_apply(const uint32_t *in, uint32_t *out, int64_t n,
                                                           const uint32_t *sums) {
    __shared__ uint32_t s_wave[4];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = base + k < n ? in[base + k] : 0u;
        s += v[k];
    }
    const uint32_t inc = block_scan_inclusive(s, s_wave, nullptr);
    uint32_t run = sums[blockIdx.x] + inc - s;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
}

int64_t scan_sums_words(int64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

int scan_exclusive_u32(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *total_host) {
    if (n <= 0) {
        if (total_host) *total_host = 0;
        return MSBG_OK;
    }
    int rc;
    uint32_t *sums = (uint32_t *)workspace(0, sizeof(uint32_t) * scan_sums_words(n), &rc);
    if (rc) return rc;
    return scan_exclusive_u32_with(in, out, n, sums, total_host);
}

int scan_exclusive_u32_with(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *sums, uint32_t *total_host) {
    if (n <= 0) {
        if (total_host) *total_host = 0;
        return MSBG_OK;
    }
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(scan_tile_sums, dim3((unsigned)tiles), dim3(kScanThreads), 0, stream(), in, n, sums);
    hipLaunchKernelGGL(scan_sums, dim3(1), dim3(kScanThreads), 0, stream(), sums, tiles, sums + tiles);
    hipLaunchKernelGGL(scan_apply, dim3((unsigned)tiles), dim3(kScanThreads), 0, stream(), in, out, n, sums);
    MSBG_HIP(hipGetLastError());
    if (total_host) {
        MSBG_HIP(hipMemcpyAsync(total_host, sums + tiles, sizeof(uint32_t), hipMemcpyDeviceToHost, stream()));
        MSBG_HIP(hipStreamSynchronize(stream()));
    }
    return MSBG_OK;
}

// ---------------------------------------------------------------------------
// Radix sort: 8-bit digits.  A tile is 256 lanes x 16 keys = 4096 keys; wave
// w of a tile owns the tile's keys [1024 w, 1024 w + 1024) as 16 chunks of 64.
// Pass = histogram (LDS atomics per tile, bin-major table), one exclusive scan
// of the table (so entry (d, tile) is the first output slot of digit d from
// that tile), then a stable scatter: inside a chunk a key's rank among equal
// digits comes from 8 ballots (peer mask), across chunks and waves from LDS
// counters, so equal keys keep their input order.
// ---------------------------------------------------------------------------
constexpr int kRsThreads = 256, kRsChunks = 16, kRsTile = kRsThreads * kRsChunks, kRsBins = 256;

__global__ __launch_bounds__(kRsThreads) void rs_histogram(const uint32_t *keys, int64_t n, int shift,
                                                           uint32_t *hist, int64_t ntiles) {
    __shared__ uint32_t h[kRsBins];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kRsTile;
#pragma unroll 4
    for (int i = threadIdx.x; i < kRsTile; i += kRsThreads)
        if (base + i < n) atomicAdd(&h[(keys[base + i] >> shift) & 0xffu], 1u);
    __syncthreads();
    hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kRsThreads) void rs_scatter(const uint32_t *__restrict__ kin,
                                                         const uint32_t *__restrict__ vin, uint32_t *__restrict__ kout,
                                                         uint32_t *__restrict__ vout, int64_t n, int shift,
                                                         const uint32_t *__restrict__ offsets, int64_t ntiles) {
    __shared__ uint32_t cnt[4][kRsBins];
    __shared__ uint32_t base[kRsBins];   // first output slot of digit d from this tile
    __shared__ uint32_t lbase[kRsBins];  // first tile-local slot of digit d
    __shared__ uint32_t sk[kRsTile], sv[kRsTile];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
#pragma unroll
    for (int w = 0; w < 4; ++w) cnt[w][t] = 0;
    base[t] = offsets[(int64_t)t * ntiles + blockIdx.x];
    __syncthreads();
    const int64_t tile0 = (int64_t)blockIdx.x * kRsTile;
    const int64_t wave0 = tile0 + (int64_t)wave * (kRsChunks * 64);
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t k[kRsChunks], v[kRsChunks], off[kRsChunks];
#pragma unroll
    for (int c = 0; c < kRsChunks; ++c) {
        const int64_t idx = wave0 + c * 64 + lane;
        const bool valid = idx < n;
        k[c] = valid ? kin[idx] : 0u;
        v[c] = valid ? vin[idx] : 0u;
        const uint32_t d = (k[c] >> shift) & 0xffu;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t before = cnt[wave][d];
        off[c] = before + (uint32_t)__popcll(peers & lt);
        // The lowest lane of each peer group advances its digit's counter.
        if (valid && (peers & lt) == 0) cnt[wave][d] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    uint32_t total;  // keys of digit t in the tile
    {  // exclusive scan over the 4 waves for digit t
        uint32_t s = 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const uint32_t x = cnt[w][t];
            cnt[w][t] = s;
            s += x;
        }
        total = s;
    }
    // exclusive scan of the digit totals across the tile: lbase
    {
        uint32_t x = total;  // inclusive scan inside each wave, then wave sums
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) lbase[wave] = x;  // wave sums, read back below
        __syncthreads();
        uint32_t wsum = 0;
        for (int w = 0; w < wave; ++w) wsum += lbase[w];
        __syncthreads();
        lbase[t] = wsum + x - total;
    }
    __syncthreads();
    // Keys into LDS in output order: digit-major, stable within a digit.
#pragma unroll
    for (int c = 0; c < kRsChunks; ++c) {
        const int64_t idx = wave0 + c * 64 + lane;
        if (idx < n) {
            const uint32_t d = (k[c] >> shift) & 0xffu;
            const uint32_t l = lbase[d] + cnt[wave][d] + off[c];
            sk[l] = k[c];
            sv[l] = v[c];
        }
    }
    __syncthreads();
    // Consecutive lanes write consecutive slots of one digit's run.
    const int valid_n = (int)min((int64_t)kRsTile, n - tile0);
#pragma unroll
    for (int c = 0; c < kRsChunks; ++c) {
        const int l = c * kRsThreads + t;
        if (l < valid_n) {
            const uint32_t kk = sk[l];
            const uint32_t d = (kk >> shift) & 0xffu;
            const uint32_t pos = base[d] + (uint32_t)l - lbase[d];
            kout[pos] = kk;
            vout[pos] = sv[l];
        }
    }
}

int radix_sort_pairs(uint32_t *keys, uint32_t *vals, int64_t n, int key_bits) {
    MSBG_CHECK(key_bits >= 1 && key_bits <= 32, MSBG_EINVAL, "radix sort: key_bits must be 1..32");
    MSBG_CHECK(n >= 0 && n < ((int64_t)1 << 31), MSBG_EINVAL, "radix sort: n must be below 2^31");
    if (n <= 1) return MSBG_OK;
    const int passes = (key_bits + 7) / 8;
    const int64_t ntiles = (n + kRsTile - 1) / kRsTile;
    int rc;
    uint32_t *tmp = (uint32_t *)workspace(1, sizeof(uint32_t) * 2 * n, &rc);
    if (rc) return rc;
    uint32_t *hist = (uint32_t *)workspace(2, sizeof(uint32_t) * kRsBins * ntiles, &rc);
    if (rc) return rc;
    uint32_t *ka = keys, *va = vals, *kb = tmp, *vb = tmp + n;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        hipLaunchKernelGGL(rs_histogram, dim3((unsigned)ntiles), dim3(kRsThreads), 0, stream(), ka, n, shift, hist,
                           ntiles);
        MSBG_HIP(hipGetLastError());
        rc = scan_exclusive_u32(hist, hist, kRsBins * ntiles, nullptr);
        if (rc) return rc;
        hipLaunchKernelGGL(rs_scatter, dim3((unsigned)ntiles), dim3(kRsThreads), 0, stream(), ka, va, kb, vb, n, shift,
                           hist, ntiles);
        MSBG_HIP(hipGetLastError());
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (ka != keys) {
        MSBG_HIP(hipMemcpyAsync(keys, ka, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, stream()));
        MSBG_HIP(hipMemcpyAsync(vals, va, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, stream()));
    }
    return MSBG_OK;
}

}  // namespace msbg

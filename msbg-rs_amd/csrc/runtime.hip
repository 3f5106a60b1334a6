// Runtime plumbing and the device-resident sparse grid (C ABI).
#include <algorithm>
#include <cstring>
#include <string>

#include "internal.hpp"
#include "primitives.hpp"

namespace msbg {

static thread_local std::string g_last_error;
static hipStream_t g_stream = nullptr;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return MSBG_EHIP;
}

hipStream_t stream() { return g_stream; }

// Fills n floats with v (a kernel, not hipMemset, so it can sit in a graph).
__global__ void fill_f32_kernel(float *p, int64_t n, float v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

int fill_f32(float *p, int64_t n, float v) {
    if (n <= 0) return MSBG_OK;
    int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_f32_kernel, dim3((unsigned)blocks), dim3(256), 0, g_stream, p, n, v);
    MSBG_HIP(hipGetLastError());
    return MSBG_OK;
}

// Block payload transfer: item i of a batch is block slots[i / N], cell i % N.
__global__ void scatter_blocks_kernel(float *slab, const int32_t *slots, const float *src, int64_t n, int64_t N) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n * N; i += (int64_t)gridDim.x * blockDim.x)
        slab[(int64_t)slots[i / N] * N + i % N] = src[i];
}

__global__ void gather_blocks_kernel(const float *slab, const int32_t *slots, float *dst, int64_t n, int64_t N) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n * N; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = slab[(int64_t)slots[i / N] * N + i % N];
}

// Sets every payload of `slots` (device list) to v.
__global__ void fill_slots_kernel(float *slab, const int32_t *slots, int64_t N, float v) {
    float *p = slab + (int64_t)slots[blockIdx.x] * N;
    for (int64_t i = threadIdx.x; i < N; i += blockDim.x) p[i] = v;
}

}  // namespace msbg

using namespace msbg;

extern "C" {

const char *msbg_last_error(void) { return g_last_error.c_str(); }

int msbg_device_count(int *n) {
    MSBG_CHECK(n, MSBG_EINVAL, "msbg_device_count: null out");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) c = 0;
    *n = c;
    return MSBG_OK;
}

int msbg_set_stream(void *s) {
    g_stream = (hipStream_t)s;
    return MSBG_OK;
}

int msbg_synchronize(void) {
    MSBG_HIP(hipStreamSynchronize(g_stream));
    return MSBG_OK;
}

int msbg_malloc(void **p, size_t bytes) {
    MSBG_CHECK(p, MSBG_EINVAL, "msbg_malloc: null out");
    *p = nullptr;
    if (bytes == 0) return MSBG_OK;
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipErrorOutOfMemory) return fail(MSBG_ENOMEM, "msbg_malloc: out of device memory");
    MSBG_HIP(e);
    return MSBG_OK;
}

int msbg_free(void *p) {
    if (p) MSBG_HIP(hipFree(p));
    return MSBG_OK;
}

int msbg_memcpy_htod(void *d, const void *h, size_t bytes) {
    if (!bytes) return MSBG_OK;
    MSBG_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, g_stream));
    MSBG_HIP(hipStreamSynchronize(g_stream));
    return MSBG_OK;
}

int msbg_memcpy_dtoh(void *h, const void *d, size_t bytes) {
    if (!bytes) return MSBG_OK;
    MSBG_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, g_stream));
    MSBG_HIP(hipStreamSynchronize(g_stream));
    return MSBG_OK;
}

int msbg_memcpy_dtod(void *d, const void *s, size_t bytes) {
    if (!bytes) return MSBG_OK;
    MSBG_HIP(hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToDevice, g_stream));
    MSBG_HIP(hipStreamSynchronize(g_stream));
    return MSBG_OK;
}

int msbg_fill_u32(void *d, int64_t n_words, uint32_t value) {
    MSBG_CHECK(n_words >= 0 && (n_words == 0 || d), MSBG_EINVAL, "msbg_fill_u32: bad buffer");
    MSBG_CHECK(((uintptr_t)d & 3u) == 0, MSBG_EINVAL, "msbg_fill_u32: the buffer must be 4-byte aligned");
    return fill_f32(reinterpret_cast<float *>(d), n_words, __builtin_bit_cast(float, value));
}

int msbg_event_create(void **ev) {
    MSBG_CHECK(ev, MSBG_EINVAL, "msbg_event_create: null out");
    hipEvent_t e;
    MSBG_HIP(hipEventCreate(&e));
    *ev = (void *)e;
    return MSBG_OK;
}

int msbg_event_destroy(void *ev) {
    if (ev) MSBG_HIP(hipEventDestroy((hipEvent_t)ev));
    return MSBG_OK;
}

int msbg_event_record(void *ev) {
    MSBG_HIP(hipEventRecord((hipEvent_t)ev, g_stream));
    return MSBG_OK;
}

int msbg_event_elapsed_ms(void *a, void *b, float *ms) {
    MSBG_HIP(hipEventSynchronize((hipEvent_t)b));
    MSBG_HIP(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
    return MSBG_OK;
}

// ---------------------------------------------------------------------------
// Sparse grid.
// ---------------------------------------------------------------------------
int msbg_grid_new(int64_t sx, int64_t sy, int64_t sz, int log2_bsx, float empty_value, float full_value,
                  int64_t capacity_blocks, msbg_grid **out) {
    MSBG_CHECK(out, MSBG_EINVAL, "msbg_grid_new: null out");
    *out = nullptr;
    MSBG_CHECK(sx > 0 && sy > 0 && sz > 0, MSBG_EINVAL, "msbg_grid_new: domain must be non-empty");
    MSBG_CHECK(log2_bsx >= 0 && log2_bsx <= 4, MSBG_EINVAL, "msbg_grid_new: block edge must be 1..16");
    MSBG_CHECK(capacity_blocks >= 0, MSBG_EINVAL, "msbg_grid_new: negative capacity");
    MSBG_CHECK(sx < (1 << 30) && sy < (1 << 30) && sz < (1 << 30), MSBG_EINVAL, "msbg_grid_new: domain too large");
    msbg_grid *g = new msbg_grid();
    g->sx = sx;
    g->sy = sy;
    g->sz = sz;
    g->L = log2_bsx;
    g->B = 1 << log2_bsx;
    g->N = (int64_t)g->B * g->B * g->B;
    g->nx = (sx + g->B - 1) >> g->L;
    g->ny = (sy + g->B - 1) >> g->L;
    g->nz = (sz + g->B - 1) >> g->L;
    g->nxy = g->nx * g->ny;
    g->n_blocks = g->nxy * g->nz;
    if (g->n_blocks > (int64_t)1 << 31) {
        delete g;
        return fail(MSBG_EINVAL, "msbg_grid_new: more than 2^31 blocks needs a sparse bid table");
    }
    g->capacity_slots = capacity_blocks + 2;
    g->n_slots = 2;
    g->empty_value = empty_value;
    g->full_value = full_value;
    g->b2s_host.assign((size_t)g->n_blocks, 0);
    hipError_t e = hipMalloc(&g->d_slab, sizeof(float) * g->capacity_slots * g->N);
    if (e == hipSuccess) e = hipMalloc(&g->d_b2s, sizeof(int32_t) * g->n_blocks);
    if (e != hipSuccess) {
        msbg_grid_free(g);
        return e == hipErrorOutOfMemory ? fail(MSBG_ENOMEM, "msbg_grid_new: out of device memory")
                                        : hip_fail(e, "msbg_grid_new");
    }
    int rc = fill_f32(g->d_slab, g->N, empty_value);
    if (rc == MSBG_OK) rc = fill_f32(g->d_slab + g->N, g->N, full_value);
    if (rc == MSBG_OK) {
        e = hipMemsetAsync(g->d_b2s, 0, sizeof(int32_t) * g->n_blocks, g_stream);
        if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
        if (e != hipSuccess) rc = hip_fail(e, "msbg_grid_new");
    }
    if (rc != MSBG_OK) {
        msbg_grid_free(g);
        return rc;
    }
    *out = g;
    return MSBG_OK;
}

int msbg_grid_free(msbg_grid *g) {
    if (!g) return MSBG_OK;
    if (g->d_slab) (void)hipFree(g->d_slab);
    if (g->d_b2s) (void)hipFree(g->d_b2s);
    delete g;
    return MSBG_OK;
}

int msbg_grid_dims(const msbg_grid *g, int64_t *d) {
    MSBG_CHECK(g && d, MSBG_EINVAL, "msbg_grid_dims: null argument");
    d[0] = g->sx;
    d[1] = g->sy;
    d[2] = g->sz;
    d[3] = g->nx;
    d[4] = g->ny;
    d[5] = g->nz;
    d[6] = g->n_blocks;
    d[7] = g->n_slots;
    return MSBG_OK;
}

static int check_bids(const msbg_grid *g, const int64_t *bids, int64_t n, const char *who) {
    MSBG_CHECK(g, MSBG_EINVAL, std::string(who) + ": null grid");
    MSBG_CHECK(n >= 0 && (n == 0 || bids), MSBG_EINVAL, std::string(who) + ": bad bid list");
    for (int64_t i = 0; i < n; ++i)
        MSBG_CHECK(bids[i] >= 0 && bids[i] < g->n_blocks, MSBG_EINVAL,
                   std::string(who) + ": bid " + std::to_string(bids[i]) + " out of range");
    return MSBG_OK;
}

static int upload_b2s(msbg_grid *g) {
    MSBG_HIP(hipMemcpyAsync(g->d_b2s, g->b2s_host.data(), sizeof(int32_t) * g->n_blocks, hipMemcpyHostToDevice,
                            g_stream));
    MSBG_HIP(hipStreamSynchronize(g_stream));
    return MSBG_OK;
}

int msbg_grid_ensure_blocks(msbg_grid *g, const int64_t *bids, int64_t n) {
    int rc = check_bids(g, bids, n, "msbg_grid_ensure_blocks");
    if (rc) return rc;
    std::vector<int64_t> fresh;
    for (int64_t i = 0; i < n; ++i)
        if (g->b2s_host[bids[i]] < 2) fresh.push_back(bids[i]);
    if (fresh.empty()) return MSBG_OK;
    std::vector<int64_t> sorted = fresh;
    std::sort(sorted.begin(), sorted.end());
    MSBG_CHECK(std::adjacent_find(sorted.begin(), sorted.end()) == sorted.end(), MSBG_EINVAL,
               "msbg_grid_ensure_blocks: bids must be unique (sparse_grid.rs:331-332)");
    MSBG_CHECK(g->n_slots + (int64_t)fresh.size() <= g->capacity_slots, MSBG_ENOMEM,
               "BlockPool memory limit exhausted (grid capacity)");
    std::vector<int32_t> slots(fresh.size());
    for (size_t i = 0; i < fresh.size(); ++i) {
        slots[i] = (int32_t)(g->n_slots + (int64_t)i);
        g->b2s_host[fresh[i]] = slots[i];
    }
    g->n_slots += (int64_t)fresh.size();
    int32_t *d_slots = nullptr;
    MSBG_HIP(hipMalloc(&d_slots, sizeof(int32_t) * slots.size()));
    hipError_t e = hipMemcpyAsync(d_slots, slots.data(), sizeof(int32_t) * slots.size(), hipMemcpyHostToDevice,
                                  g_stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(fill_slots_kernel, dim3((unsigned)slots.size()), dim3(256), 0, g_stream, g->d_slab,
                           d_slots, g->N, g->empty_value);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
    (void)hipFree(d_slots);
    MSBG_HIP(e);
    return upload_b2s(g);
}

int msbg_grid_fill_blocks(msbg_grid *g, const int64_t *bids, int64_t n, float value) {
    int rc = check_bids(g, bids, n, "msbg_grid_fill_blocks");
    if (rc) return rc;
    std::vector<int32_t> slots;
    for (int64_t i = 0; i < n; ++i)
        if (g->b2s_host[bids[i]] >= 2) slots.push_back(g->b2s_host[bids[i]]);
    if (slots.empty()) return MSBG_OK;
    int32_t *d_slots = nullptr;
    MSBG_HIP(hipMalloc(&d_slots, sizeof(int32_t) * slots.size()));
    hipError_t e = hipMemcpyAsync(d_slots, slots.data(), sizeof(int32_t) * slots.size(), hipMemcpyHostToDevice,
                                  g_stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(fill_slots_kernel, dim3((unsigned)slots.size()), dim3(256), 0, g_stream, g->d_slab,
                           d_slots, g->N, value);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
    (void)hipFree(d_slots);
    MSBG_HIP(e);
    return MSBG_OK;
}

int msbg_grid_set_full_block(msbg_grid *g, int64_t bid) {
    int rc = check_bids(g, &bid, 1, "msbg_grid_set_full_block");
    if (rc) return rc;
    MSBG_CHECK(g->b2s_host[bid] < 2, MSBG_EINVAL, "set_full_block on an allocated block (sparse_grid.rs:280)");
    g->b2s_host[bid] = 1;
    return upload_b2s(g);
}

int msbg_grid_set_empty_block(msbg_grid *g, int64_t bid) {
    int rc = check_bids(g, &bid, 1, "msbg_grid_set_empty_block");
    if (rc) return rc;
    MSBG_CHECK(g->b2s_host[bid] < 2, MSBG_EINVAL, "set_empty_block on an allocated block (sparse_grid.rs:272)");
    g->b2s_host[bid] = 0;
    return upload_b2s(g);
}

int msbg_grid_set_empty_value(msbg_grid *g, float v) {
    MSBG_CHECK(g, MSBG_EINVAL, "null grid");
    g->empty_value = v;
    int rc = fill_f32(g->d_slab, g->N, v);
    if (rc) return rc;
    MSBG_HIP(hipStreamSynchronize(g_stream));
    return MSBG_OK;
}

int msbg_grid_set_full_value(msbg_grid *g, float v) {
    MSBG_CHECK(g, MSBG_EINVAL, "null grid");
    g->full_value = v;
    int rc = fill_f32(g->d_slab + g->N, g->N, v);
    if (rc) return rc;
    MSBG_HIP(hipStreamSynchronize(g_stream));
    return MSBG_OK;
}

int msbg_grid_write_blocks(msbg_grid *g, const int64_t *bids, int64_t n, const float *host) {
    int rc = check_bids(g, bids, n, "msbg_grid_write_blocks");
    if (rc) return rc;
    MSBG_CHECK(n == 0 || host, MSBG_EINVAL, "msbg_grid_write_blocks: null data");
    std::vector<int32_t> slots((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        MSBG_CHECK(g->b2s_host[bids[i]] >= 2, MSBG_EINVAL, "msbg_grid_write_blocks: block is not allocated");
        slots[(size_t)i] = g->b2s_host[bids[i]];
    }
    if (n == 0) return MSBG_OK;
    // one staged upload of the payload and the slot list, then a scatter kernel
    float *stage = (float *)workspace(20, sizeof(float) * n * g->N + sizeof(int32_t) * n, &rc);
    if (rc) return rc;
    int32_t *d_slots = (int32_t *)(stage + n * g->N);
    MSBG_HIP(hipMemcpyAsync(stage, host, sizeof(float) * n * g->N, hipMemcpyHostToDevice, g_stream));
    MSBG_HIP(hipMemcpyAsync(d_slots, slots.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, g_stream));
    hipLaunchKernelGGL(scatter_blocks_kernel, dim3((unsigned)std::min<int64_t>((n * g->N + 255) / 256, 65536)),
                       dim3(256), 0, g_stream, g->d_slab, d_slots, stage, n, g->N);
    MSBG_HIP(hipGetLastError());
    MSBG_HIP(hipStreamSynchronize(g_stream));
    return MSBG_OK;
}

int msbg_grid_read_blocks(const msbg_grid *g, const int64_t *bids, int64_t n, float *host) {
    int rc = check_bids(g, bids, n, "msbg_grid_read_blocks");
    if (rc) return rc;
    MSBG_CHECK(n == 0 || host, MSBG_EINVAL, "msbg_grid_read_blocks: null out");
    if (n == 0) return MSBG_OK;
    std::vector<int32_t> slots((size_t)n);
    for (int64_t i = 0; i < n; ++i) slots[(size_t)i] = g->b2s_host[bids[i]];
    // a gather kernel into staging, then one download
    float *stage = (float *)workspace(20, sizeof(float) * n * g->N + sizeof(int32_t) * n, &rc);
    if (rc) return rc;
    int32_t *d_slots = (int32_t *)(stage + n * g->N);
    MSBG_HIP(hipMemcpyAsync(d_slots, slots.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, g_stream));
    hipLaunchKernelGGL(gather_blocks_kernel, dim3((unsigned)std::min<int64_t>((n * g->N + 255) / 256, 65536)),
                       dim3(256), 0, g_stream, g->d_slab, d_slots, stage, n, g->N);
    MSBG_HIP(hipGetLastError());
    MSBG_HIP(hipMemcpyAsync(host, stage, sizeof(float) * n * g->N, hipMemcpyDeviceToHost, g_stream));
    MSBG_HIP(hipStreamSynchronize(g_stream));
    return MSBG_OK;
}

int msbg_grid_load(msbg_grid *g, const int32_t *b2s, const float *slab, int64_t n_slots) {
    MSBG_CHECK(g && b2s && slab, MSBG_EINVAL, "msbg_grid_load: null argument");
    MSBG_CHECK(n_slots >= 2 && n_slots <= g->capacity_slots, MSBG_ENOMEM, "msbg_grid_load: slots exceed capacity");
    for (int64_t b = 0; b < g->n_blocks; ++b)
        MSBG_CHECK(b2s[b] >= 0 && b2s[b] < n_slots, MSBG_EINVAL, "msbg_grid_load: b2s entry out of range");
    g->b2s_host.assign(b2s, b2s + g->n_blocks);
    g->n_slots = n_slots;
    g->empty_value = slab[0];
    g->full_value = slab[g->N];
    MSBG_HIP(hipMemcpyAsync(g->d_slab, slab, sizeof(float) * n_slots * g->N, hipMemcpyHostToDevice, g_stream));
    return upload_b2s(g);
}

int msbg_grid_store(const msbg_grid *g, int32_t *b2s, float *slab, int64_t cap_slots, int64_t *n_slots) {
    MSBG_CHECK(g && b2s && slab && n_slots, MSBG_EINVAL, "msbg_grid_store: null argument");
    MSBG_CHECK(cap_slots >= g->n_slots, MSBG_EINVAL, "msbg_grid_store: output slab too small");
    std::memcpy(b2s, g->b2s_host.data(), sizeof(int32_t) * g->n_blocks);
    MSBG_HIP(hipMemcpyAsync(slab, g->d_slab, sizeof(float) * g->n_slots * g->N, hipMemcpyDeviceToHost, g_stream));
    MSBG_HIP(hipStreamSynchronize(g_stream));
    *n_slots = g->n_slots;
    return MSBG_OK;
}

int msbg_grid_copy(msbg_grid *dst, const msbg_grid *src) {
    MSBG_CHECK(dst && src, MSBG_EINVAL, "msbg_grid_copy: null argument");
    MSBG_CHECK(dst->sx == src->sx && dst->sy == src->sy && dst->sz == src->sz && dst->L == src->L, MSBG_EINVAL,
               "msbg_grid_copy: grids must share dimensions and block size");
    MSBG_CHECK(dst->capacity_slots >= src->n_slots, MSBG_ENOMEM, "msbg_grid_copy: destination capacity too small");
    if (dst == src) return MSBG_OK;
    MSBG_HIP(hipMemcpyAsync(dst->d_b2s, src->d_b2s, sizeof(int32_t) * src->n_blocks, hipMemcpyDeviceToDevice,
                            g_stream));
    MSBG_HIP(hipMemcpyAsync(dst->d_slab, src->d_slab, sizeof(float) * src->n_slots * src->N,
                            hipMemcpyDeviceToDevice, g_stream));
    dst->b2s_host = src->b2s_host;
    dst->n_slots = src->n_slots;
    dst->empty_value = src->empty_value;
    dst->full_value = src->full_value;
    return MSBG_OK;
}

int msbg_grid_block_slots(const msbg_grid *g, int32_t *b2s) {
    MSBG_CHECK(g && b2s, MSBG_EINVAL, "msbg_grid_block_slots: null argument");
    std::memcpy(b2s, g->b2s_host.data(), sizeof(int32_t) * g->n_blocks);
    return MSBG_OK;
}

int msbg_grid_value_block_ptr(msbg_grid *g, int64_t bid, float **p) {
    int rc = check_bids(g, &bid, 1, "msbg_grid_value_block_ptr");
    if (rc) return rc;
    MSBG_CHECK(p, MSBG_EINVAL, "null out");
    int32_t s = g->b2s_host[bid];
    *p = s >= 2 ? g->d_slab + (int64_t)s * g->N : nullptr;
    return MSBG_OK;
}

int msbg_grid_device_views(msbg_grid *g, float **slab, int32_t **b2s) {
    MSBG_CHECK(g, MSBG_EINVAL, "null grid");
    if (slab) *slab = g->d_slab;
    if (b2s) *b2s = g->d_b2s;
    return MSBG_OK;
}

}  // extern "C"

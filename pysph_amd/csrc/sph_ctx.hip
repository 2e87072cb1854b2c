// Context, error handling, scratch management and device-wide scans.
#include <stdlib.h>
#include <string.h>

#include "sph_internal.h"

extern "C" {

const char *sph_version(void) { return SPH_VERSION_STRING; }

int sph_ctx_create(sph_ctx **out, int device, void *stream)
{
    if (!out) return SPH_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return SPH_EHIP;
    if (device < 0 || device >= ndev) return SPH_EINVAL;
    sph_ctx *ctx = (sph_ctx *)calloc(1, sizeof(sph_ctx));
    if (!ctx) return SPH_ENOMEM;
    ctx->device = device;
    ctx->stream = (hipStream_t)stream;
    ctx->key_order = -1;
    if (hipSetDevice(device) != hipSuccess) {
        free(ctx);
        return SPH_EHIP;
    }
    // the scratch address of sph_extrema_noalloc; on failure that entry
    // point tries again on its first call and reports the error
    (void)sph_extrema_scratch_setup(ctx);
    *out = ctx;
    return SPH_OK;
}

int sph_ctx_destroy(sph_ctx *ctx)
{
    if (!ctx) return SPH_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    else (void)hipDeviceSynchronize();
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->key_tab) (void)hipFree(ctx->key_tab);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    free(ctx);
    return SPH_OK;
}

int sph_ctx_set_stream(sph_ctx *ctx, void *stream)
{
    if (!ctx) return SPH_EINVAL;
    ctx->stream = (hipStream_t)stream;
    return SPH_OK;
}

int sph_ctx_synchronize(sph_ctx *ctx)
{
    if (!ctx) return SPH_EINVAL;
    SPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return SPH_OK;
}

const char *sph_last_error(sph_ctx *ctx)
{
    if (!ctx) return "null sph_ctx";
    return ctx->err;
}

}  // extern "C"

int sph_set_error(sph_ctx *ctx, int code, const char *fmt, ...)
{
    if (ctx) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(ctx->err, sizeof(ctx->err), fmt, ap);
        va_end(ap);
    }
    return code;
}

int sph_ensure_scratch(sph_ctx *ctx, size_t bytes)
{
    if (bytes <= ctx->scratch_bytes) return SPH_OK;
    if (ctx->frozen)
        return sph_set_error(ctx, SPH_EINVAL,
                             "scratch of %zu bytes needed but a captured graph holds the "
                             "current one (%zu bytes)", bytes, ctx->scratch_bytes);
    size_t nb = bytes + bytes / 4 + 4096;
    if (ctx->scratch) {
        SPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        SPH_HIP_CHECK(ctx, hipFree(ctx->scratch));
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
    }
    if (hipMalloc(&ctx->scratch, nb) != hipSuccess) {
        ctx->scratch = nullptr;
        return sph_set_error(ctx, SPH_ENOMEM, "scratch allocation of %zu bytes failed", nb);
    }
    ctx->scratch_bytes = nb;
    return SPH_OK;
}

int sph_cu_count(sph_ctx *ctx)
{
    if (ctx->ncu <= 0) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, ctx->device) !=
                hipSuccess ||
            v <= 0)
            v = 256;  // MI355X
        ctx->ncu = v;
    }
    return ctx->ncu;
}

int sph_ensure_pinned(sph_ctx *ctx, size_t bytes)
{
    if (bytes <= ctx->pinned_bytes) return SPH_OK;
    if (ctx->pinned) {
        SPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        SPH_HIP_CHECK(ctx, hipHostFree(ctx->pinned));
        ctx->pinned = nullptr;
        ctx->pinned_bytes = 0;
    }
    size_t nb = bytes < 4096 ? 4096 : bytes;
    if (hipHostMalloc(&ctx->pinned, nb, hipHostMallocDefault) != hipSuccess) {
        ctx->pinned = nullptr;
        return sph_set_error(ctx, SPH_ENOMEM, "pinned allocation of %zu bytes failed", nb);
    }
    ctx->pinned_bytes = nb;
    return SPH_OK;
}

// ---------------------------------------------------------------------------
// Device-wide exclusive scan.  Three phases over tiles of SCAN_TILE elements:
// tile sums -> (recursive) scan of the tile sums -> tile scan plus offset.
// Inside a tile every thread scans SCAN_ITEMS consecutive values serially,
// then the 64-lane wavefronts scan the thread totals with shuffles and the
// wavefront totals are combined through LDS.
// ---------------------------------------------------------------------------
#define SCAN_THREADS 256
#define SCAN_ITEMS 8
#define SCAN_TILE (SCAN_THREADS * SCAN_ITEMS)

template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// exclusive scan of one value per thread across the block; returns the
// exclusive prefix and the block total in *total
template <typename T>
__device__ __forceinline__ T block_exclusive_scan(T v, T *total)
{
    __shared__ T wsum[SCAN_THREADS / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    T inc = wave_inclusive_scan(v);
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    T woff = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < SCAN_THREADS / 64; ++w) {
        T s = wsum[w];
        if (w < wid) woff += s;
        tot += s;
    }
    *total = tot;
    __syncthreads();
    return woff + inc - v;
}

template <typename T>
__global__ void __launch_bounds__(SCAN_THREADS) k_tile_sums(const T *__restrict__ data, int64_t n,
                                                            T *__restrict__ sums)
{
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    T s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k)
        if (base + k < n) s += data[base + k];
    T total;
    block_exclusive_scan(s, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

template <typename T>
__global__ void __launch_bounds__(SCAN_THREADS) k_tile_scan(T *__restrict__ data, int64_t n,
                                                            const T *__restrict__ offs,
                                                            T *__restrict__ total_out)
{
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    T v[SCAN_ITEMS];
    T s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        v[k] = (base + k < n) ? data[base + k] : (T)0;
        s += v[k];
    }
    T total;
    T pre = block_exclusive_scan(s, &total);
    T run = pre + (offs ? offs[blockIdx.x] : (T)0);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        if (base + k < n) data[base + k] = run;
        run += v[k];
    }
    if (total_out && blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_THREADS - 1)
        *total_out = run;
}

static int scan_levels(int64_t n, size_t elem, int64_t *levels_n, size_t *levels_off,
                       size_t *bytes_out)
{
    int nl = 0;
    size_t bytes = 0;
    int64_t m = n;
    while (m > SCAN_TILE) {
        if (nl >= 7) return -1;
        int64_t nb = (m + SCAN_TILE - 1) / SCAN_TILE;
        levels_n[nl] = nb;
        levels_off[nl] = bytes;
        bytes += (size_t)(nb + 1) * elem;
        bytes = (bytes + 255) & ~(size_t)255;
        ++nl;
        m = nb;
    }
    *bytes_out = bytes + 256;
    return nl;
}

size_t sph_scan_scratch_bytes(int64_t n, size_t elem)
{
    int64_t ln[8];
    size_t lo[8], bytes = 256;
    if (n > 0) scan_levels(n, elem, ln, lo, &bytes);
    return bytes;
}

// data[0] = 0 by a kernel, not hipMemsetAsync: the scans may be captured in
// a HIP graph
template <typename T> __global__ void k_zero_first(T *data)
{
    if (threadIdx.x == 0) data[0] = 0;
}

template <typename T>
static int scan_exclusive(sph_ctx *ctx, T *data, int64_t n, size_t scratch_off)
{
    if (n <= 0) {
        hipLaunchKernelGGL(k_zero_first<T>, dim3(1), dim3(64), 0, ctx->stream, data);
        SPH_LAUNCH_CHECK(ctx, "k_zero_first");
        return SPH_OK;
    }
    // tile-sum levels live in scratch (from scratch_off): n/TILE, n/TILE^2, ...
    int64_t levels_n[8];
    size_t levels_off[8];
    size_t bytes = 0;
    const int nl = scan_levels(n, sizeof(T), levels_n, levels_off, &bytes);
    if (nl < 0) return sph_set_error(ctx, SPH_EINVAL, "scan too large");
    scratch_off = (scratch_off + 255) & ~(size_t)255;
    int rc = sph_ensure_scratch(ctx, scratch_off + bytes);
    if (rc) return rc;
    char *base = (char *)ctx->scratch + scratch_off;
    // up-sweep
    T *cur = data;
    int64_t cur_n = n;
    for (int l = 0; l < nl; ++l) {
        T *sums = (T *)(base + levels_off[l]);
        hipLaunchKernelGGL(k_tile_sums<T>, dim3((unsigned)levels_n[l]), dim3(SCAN_THREADS), 0,
                           ctx->stream, cur, cur_n, sums);
        SPH_LAUNCH_CHECK(ctx, "k_tile_sums");
        cur = sums;
        cur_n = levels_n[l];
    }
    // top level: one tile
    hipLaunchKernelGGL(k_tile_scan<T>, dim3(1), dim3(SCAN_THREADS), 0, ctx->stream, cur, cur_n,
                       (const T *)nullptr, cur + cur_n);
    SPH_LAUNCH_CHECK(ctx, "k_tile_scan");
    // down-sweep
    for (int l = nl - 1; l >= 0; --l) {
        T *lower = (l == 0) ? data : (T *)(base + levels_off[l - 1]);
        int64_t lower_n = (l == 0) ? n : levels_n[l - 1];
        T *sums = (T *)(base + levels_off[l]);
        hipLaunchKernelGGL(k_tile_scan<T>, dim3((unsigned)levels_n[l]), dim3(SCAN_THREADS), 0,
                           ctx->stream, lower, lower_n, (const T *)sums, lower + lower_n);
        SPH_LAUNCH_CHECK(ctx, "k_tile_scan");
    }
    return SPH_OK;
}

int sph_scan_exclusive_i32_at(sph_ctx *ctx, int32_t *data, int64_t n, size_t scratch_off)
{
    return scan_exclusive<int32_t>(ctx, data, n, scratch_off);
}

int sph_scan_exclusive_i64_at(sph_ctx *ctx, int64_t *data, int64_t n, size_t scratch_off)
{
    return scan_exclusive<int64_t>(ctx, data, n, scratch_off);
}

int sph_scan_exclusive_i32(sph_ctx *ctx, int32_t *data, int64_t n)
{
    return scan_exclusive<int32_t>(ctx, data, n, 0);
}

int sph_scan_exclusive_i64(sph_ctx *ctx, int64_t *data, int64_t n)
{
    return scan_exclusive<int64_t>(ctx, data, n, 0);
}

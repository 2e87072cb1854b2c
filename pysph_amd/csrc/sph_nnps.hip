// Cell binning (counting sort into Morton- or linearly-ordered cells) and
// per-destination neighbour lists in the reference's order.
//
// Reference behaviour restated (base/cell.pyx, base/nnps.pyx):
//   cell id      (int)floor(x / cell_size) per axis          cell.pyx:44-84
//   cell size    radius_scale*max(h), clamped, 0 -> 1.0       cell.pyx:911-941
//   cell content ascending particle index                      cell.pyx:1229-1306
//   stencil      x outer, y, z inner over (2d+1)^3 cells,
//                d = (int)ceil(radius/cell_size)              cell.pyx:95-118, :996-1033
//   predicate    distance2(src,dst) < (h_dst*radius_scale)^2   nnps.pyx:271-356, :552
//   self slot    own index swapped into the last slot          nnps.pyx:612-619
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sph_internal.h"

#define NN_THREADS 256

// ---------------------------------------------------------------------------
// bounding box + max/min h
// ---------------------------------------------------------------------------
// per-block partials: [xmin, ymin, zmin, -xmax, -ymax, -zmax, -hmax, hmin, nonfinite]
#define BB_VALS 9

__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fmin(v, __shfl_xor(v, d, 64));
    return v;
}

__global__ void __launch_bounds__(NN_THREADS)
    k_bbox(const double *__restrict__ x, const double *__restrict__ y,
           const double *__restrict__ z, const double *__restrict__ h, int64_t n,
           double *__restrict__ partial)
{
    double v[BB_VALS];
    v[0] = v[1] = v[2] = v[3] = v[4] = v[5] = v[6] = v[7] = INFINITY;
    double bad = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * NN_THREADS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * NN_THREADS) {
        double a = x[i], b = y[i], c = z[i], hh = h[i];
        if (!isfinite(a) || !isfinite(b) || !isfinite(c) || !isfinite(hh)) {
            bad += 1.0;
            continue;
        }
        v[0] = fmin(v[0], a);
        v[1] = fmin(v[1], b);
        v[2] = fmin(v[2], c);
        v[3] = fmin(v[3], -a);
        v[4] = fmin(v[4], -b);
        v[5] = fmin(v[5], -c);
        v[6] = fmin(v[6], -hh);
        v[7] = fmin(v[7], hh);
    }
    __shared__ double red[NN_THREADS / 64][BB_VALS];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = wave_min(v[k]);
    // count of non-finite values: reduce as a min of negatives
    double nb = -wave_min(-bad);  // max per wave, summed below is not needed: any > 0 flags
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 8; ++k) red[wid][k] = v[k];
        red[wid][8] = nb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double o[BB_VALS];
        for (int k = 0; k < BB_VALS; ++k) o[k] = red[0][k];
        for (int w = 1; w < NN_THREADS / 64; ++w) {
            for (int k = 0; k < 8; ++k) o[k] = fmin(o[k], red[w][k]);
            o[8] = fmax(o[8], red[w][8]);
        }
        for (int k = 0; k < BB_VALS; ++k) partial[(int64_t)blockIdx.x * BB_VALS + k] = o[k];
    }
}

__global__ void __launch_bounds__(NN_THREADS)
    k_bbox_final(const double *__restrict__ partial, int nparts, double *__restrict__ out)
{
    double v[BB_VALS];
    for (int k = 0; k < 8; ++k) v[k] = INFINITY;
    v[8] = 0.0;
    for (int p = threadIdx.x; p < nparts; p += NN_THREADS) {
        for (int k = 0; k < 8; ++k) v[k] = fmin(v[k], partial[(int64_t)p * BB_VALS + k]);
        v[8] = fmax(v[8], partial[(int64_t)p * BB_VALS + 8]);
    }
    __shared__ double red[NN_THREADS / 64][BB_VALS];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int k = 0; k < 8; ++k) v[k] = wave_min(v[k]);
    v[8] = -wave_min(-v[8]);
    if (lane == 0)
        for (int k = 0; k < BB_VALS; ++k) red[wid][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < NN_THREADS / 64; ++w) {
            for (int k = 0; k < 8; ++k) red[0][k] = fmin(red[0][k], red[w][k]);
            red[0][8] = fmax(red[0][8], red[w][8]);
        }
        for (int k = 0; k < BB_VALS; ++k) out[k] = red[0][k];
    }
}

// host restatement of CellManager.compute_cell_size (cell.pyx:911-941)
static double compute_cell_size(double max_h, double radius_scale, double min_size,
                                double max_size)
{
    double cell_size = radius_scale * max_h;
    if (cell_size < min_size)
        cell_size = min_size;
    else if (cell_size > max_size && max_size > 0)
        cell_size = max_size;
    if (cell_size == 0.0) cell_size = 1.0;
    return cell_size;
}

static int ceil_log2(int64_t v)
{
    int b = 0;
    while (((int64_t)1 << b) < v) ++b;
    return b;
}

// Per-axis key tables: key(cx,cy,cz) = T[cx] + T[dx + cy] + T[dx + dy + cz].
// Morton: bit l of axis a goes to position (interleave round-robin over the
// axes that still have bits at level l, z first, then y, then x), so x is
// the most significant within a level.  Linear: (cx*dy + cy)*dz + cz.
static void fill_key_tables(const sph_grid *g, int32_t *tab)
{
    const int32_t *d = g->dims;
    int32_t *tx = tab, *ty = tab + d[0], *tz = tab + d[0] + d[1];
    if (g->order == SPH_ORDER_LINEAR) {
        for (int c = 0; c < d[0]; ++c) tx[c] = (int32_t)((int64_t)c * d[1] * d[2]);
        for (int c = 0; c < d[1]; ++c) ty[c] = (int32_t)((int64_t)c * d[2]);
        for (int c = 0; c < d[2]; ++c) tz[c] = c;
        return;
    }
    int pos[3][32];
    int p = 0, maxb = g->bits[0];
    if (g->bits[1] > maxb) maxb = g->bits[1];
    if (g->bits[2] > maxb) maxb = g->bits[2];
    for (int l = 0; l < maxb; ++l)
        for (int a = 2; a >= 0; --a)
            if (l < g->bits[a]) pos[a][l] = p++;
    int32_t *t[3] = {tx, ty, tz};
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < d[a]; ++c) {
            int64_t k = 0;
            for (int l = 0; l < g->bits[a]; ++l)
                if ((c >> l) & 1) k |= (int64_t)1 << pos[a][l];
            t[a][c] = (int32_t)k;
        }
}

int sph_upload_key_tables(sph_ctx *ctx, const sph_grid *g)
{
    if (ctx->key_tab && ctx->key_order == g->order && ctx->key_dims[0] == g->dims[0] &&
        ctx->key_dims[1] == g->dims[1] && ctx->key_dims[2] == g->dims[2] &&
        ctx->key_bits[0] == g->bits[0] && ctx->key_bits[1] == g->bits[1] &&
        ctx->key_bits[2] == g->bits[2])
        return SPH_OK;
    int64_t len = (int64_t)g->dims[0] + g->dims[1] + g->dims[2];
    if (len > ctx->key_tab_len) {
        if (ctx->key_tab) {
            SPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
            SPH_HIP_CHECK(ctx, hipFree(ctx->key_tab));
            ctx->key_tab = nullptr;
        }
        SPH_HIP_CHECK(ctx, hipMalloc(&ctx->key_tab, sizeof(int32_t) * (size_t)len));
        ctx->key_tab_len = len;
    }
    int32_t *h = (int32_t *)malloc(sizeof(int32_t) * (size_t)len);
    if (!h) return sph_set_error(ctx, SPH_ENOMEM, "host key table");
    fill_key_tables(g, h);
    // stream-ordered upload through pinned staging so the copy is async-safe
    int rc = sph_ensure_pinned(ctx, sizeof(int32_t) * (size_t)len);
    if (rc) {
        free(h);
        return rc;
    }
    SPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    memcpy(ctx->pinned, h, sizeof(int32_t) * (size_t)len);
    free(h);
    SPH_HIP_CHECK(ctx, hipMemcpyAsync(ctx->key_tab, ctx->pinned, sizeof(int32_t) * (size_t)len,
                                      hipMemcpyHostToDevice, ctx->stream));
    SPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->key_order = g->order;
    for (int a = 0; a < 3; ++a) {
        ctx->key_dims[a] = g->dims[a];
        ctx->key_bits[a] = g->bits[a];
    }
    return SPH_OK;
}

extern "C" int sph_grid_build(sph_ctx *ctx, int narrays, const double *const *xs,
                              const double *const *ys, const double *const *zs,
                              const double *const *hs, const int64_t *ns, double radius_scale,
                              double min_cell_size, double max_cell_size, int order,
                              int64_t max_cells, sph_grid *grid)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, grid && narrays >= 0 && xs && ys && zs && hs && ns, "sph_grid_build: null");
    SPH_REQUIRE(ctx, order == SPH_ORDER_MORTON || order == SPH_ORDER_LINEAR,
                "sph_grid_build: bad order %d", order);
    const int kBlocks = 512;
    int nparts = 0;
    for (int a = 0; a < narrays; ++a)
        if (ns[a] > 0) nparts += kBlocks;
    memset(grid, 0, sizeof(*grid));
    grid->order = order;
    if (nparts == 0) {  // no particles: CellManager default cell size 1.0 ("h == 0" hack)
        grid->cell_size = compute_cell_size(-1.0, radius_scale, min_cell_size, max_cell_size);
        grid->max_h = -1.0;
        grid->min_h = 1e100;
        for (int a = 0; a < 3; ++a) grid->dims[a] = 1;
        grid->ncells = 1;
        return sph_upload_key_tables(ctx, grid);
    }
    size_t bytes = sizeof(double) * BB_VALS * (size_t)(nparts + 1);
    int rc = sph_ensure_scratch(ctx, bytes);
    if (rc) return rc;
    double *partial = (double *)ctx->scratch;
    int off = 0;
    for (int a = 0; a < narrays; ++a) {
        if (ns[a] <= 0) continue;
        hipLaunchKernelGGL(k_bbox, dim3(kBlocks), dim3(NN_THREADS), 0, ctx->stream, xs[a],
                           ys[a], zs[a], hs[a], ns[a], partial + (size_t)off * BB_VALS);
        SPH_LAUNCH_CHECK(ctx, "k_bbox");
        off += kBlocks;
    }
    double *fin = partial + (size_t)nparts * BB_VALS;
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(NN_THREADS), 0, ctx->stream, partial, nparts,
                       fin);
    SPH_LAUNCH_CHECK(ctx, "k_bbox_final");
    rc = sph_ensure_pinned(ctx, sizeof(double) * BB_VALS);
    if (rc) return rc;
    SPH_HIP_CHECK(ctx, hipMemcpyAsync(ctx->pinned, fin, sizeof(double) * BB_VALS,
                                      hipMemcpyDeviceToHost, ctx->stream));
    SPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    double v[BB_VALS];
    memcpy(v, ctx->pinned, sizeof(v));
    if (v[8] > 0.0)
        return sph_set_error(ctx, SPH_EDOMAIN, "non-finite particle coordinate or h");
    double max_h = -v[6];
    grid->max_h = max_h;
    grid->min_h = v[7];
    double cs = compute_cell_size(max_h, radius_scale, min_cell_size, max_cell_size);
    if (!(cs > 0.0) || !isfinite(cs))
        return sph_set_error(ctx, SPH_EDOMAIN, "bad cell size %g", cs);
    grid->cell_size = cs;
    double lo[3] = {v[0], v[1], v[2]}, hi[3] = {-v[3], -v[4], -v[5]};
    int64_t total = 1;
    for (int a = 0; a < 3; ++a) {
        double fl = floor(lo[a] / cs), fh = floor(hi[a] / cs);
        if (fabs(fl) > 1.0e9 || fabs(fh) > 1.0e9)
            return sph_set_error(ctx, SPH_EDOMAIN, "cell ids out of range on axis %d", a);
        grid->cmin[a] = (int32_t)fl;
        int64_t d = (int64_t)fh - (int64_t)fl + 1;
        if (d > (1 << 21))
            return sph_set_error(ctx, SPH_EDOMAIN, "grid too large on axis %d (%lld cells)", a,
                                 (long long)d);
        grid->dims[a] = (int32_t)d;
        grid->bits[a] = ceil_log2(d);
    }
    if (order == SPH_ORDER_MORTON)
        total = (int64_t)1 << (grid->bits[0] + grid->bits[1] + grid->bits[2]);
    else
        total = (int64_t)grid->dims[0] * grid->dims[1] * grid->dims[2];
    if (max_cells <= 0 || max_cells > ((int64_t)1 << 31) - 2) max_cells = ((int64_t)1 << 31) - 2;
    if (total > max_cells)
        return sph_set_error(ctx, SPH_EDOMAIN,
                             "cell table of %lld entries exceeds the limit %lld "
                             "(dims %d x %d x %d)",
                             (long long)total, (long long)max_cells, grid->dims[0],
                             grid->dims[1], grid->dims[2]);
    grid->ncells = total;
    return sph_upload_key_tables(ctx, grid);
}

// ---------------------------------------------------------------------------
// binning
// ---------------------------------------------------------------------------
struct GridDev {
    double cs;
    int cmin0, cmin1, cmin2;
    int d0, d1, d2;
    const int32_t *tab;  // tx | ty | tz
};

static GridDev make_griddev(sph_ctx *ctx, const sph_grid *g)
{
    GridDev d;
    d.cs = g->cell_size;
    d.cmin0 = g->cmin[0];
    d.cmin1 = g->cmin[1];
    d.cmin2 = g->cmin[2];
    d.d0 = g->dims[0];
    d.d1 = g->dims[1];
    d.d2 = g->dims[2];
    d.tab = ctx->key_tab;
    return d;
}

__device__ __forceinline__ int cell_coord(double v, double cs) { return (int)floor(v / cs); }

__device__ __forceinline__ int32_t cell_key(const GridDev &g, int cx, int cy, int cz)
{
    return g.tab[cx] + g.tab[g.d0 + cy] + g.tab[g.d0 + g.d1 + cz];
}

__global__ void k_fill_i32(int32_t *__restrict__ p, int64_t n, int32_t v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void __launch_bounds__(NN_THREADS)
    k_bin_keys(GridDev g, const double *__restrict__ x, const double *__restrict__ y,
               const double *__restrict__ z, int64_t n, int32_t *__restrict__ count,
               int32_t *__restrict__ cell_of, int32_t *__restrict__ rank)
{
    int64_t i = (int64_t)blockIdx.x * NN_THREADS + threadIdx.x;
    if (i >= n) return;
    int cx = cell_coord(x[i], g.cs) - g.cmin0;
    int cy = cell_coord(y[i], g.cs) - g.cmin1;
    int cz = cell_coord(z[i], g.cs) - g.cmin2;
    // every particle lies in the grid by construction (sph_grid_build); clamp
    // defensively so a stale grid can never index outside the tables
    cx = min(max(cx, 0), g.d0 - 1);
    cy = min(max(cy, 0), g.d1 - 1);
    cz = min(max(cz, 0), g.d2 - 1);
    int32_t key = cell_key(g, cx, cy, cz);
    cell_of[i] = key;
    rank[i] = atomicAdd(&count[key], 1);
}

__global__ void __launch_bounds__(NN_THREADS)
    k_bin_scatter(const int32_t *__restrict__ cell_of, const int32_t *__restrict__ rank,
                  int64_t n, const int32_t *__restrict__ start, int32_t *__restrict__ perm)
{
    int64_t i = (int64_t)blockIdx.x * NN_THREADS + threadIdx.x;
    if (i >= n) return;
    perm[start[cell_of[i]] + rank[i]] = (int32_t)i;
}

// Restore ascending particle index inside every cell (the atomics above
// hand out ranks in arbitrary order).  One thread per cell; cells hold a
// handful of particles, insertion sort is the right tool.
__global__ void __launch_bounds__(NN_THREADS)
    k_cell_sort(const int32_t *__restrict__ start, int64_t ncells, int32_t *__restrict__ perm)
{
    int64_t c = (int64_t)blockIdx.x * NN_THREADS + threadIdx.x;
    if (c >= ncells) return;
    int32_t s = start[c], e = start[c + 1];
    for (int32_t a = s + 1; a < e; ++a) {
        int32_t v = perm[a];
        int32_t b = a - 1;
        while (b >= s && perm[b] > v) {
            perm[b + 1] = perm[b];
            --b;
        }
        perm[b + 1] = v;
    }
}

__global__ void __launch_bounds__(NN_THREADS)
    k_gather3(const int32_t *__restrict__ perm, int64_t n, const double *__restrict__ x,
              const double *__restrict__ y, const double *__restrict__ z,
              double *__restrict__ sx, double *__restrict__ sy, double *__restrict__ sz)
{
    int64_t p = (int64_t)blockIdx.x * NN_THREADS + threadIdx.x;
    if (p >= n) return;
    int32_t i = perm[p];
    sx[p] = x[i];
    sy[p] = y[i];
    sz[p] = z[i];
}

extern "C" int sph_bin_array(sph_ctx *ctx, const sph_grid *grid, const double *x,
                             const double *y, const double *z, int64_t n, int32_t *cell_start,
                             int32_t *perm, double *sx, double *sy, double *sz,
                             int32_t *cell_of)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, grid && grid->ncells > 0 && cell_start, "sph_bin_array: bad grid");
    SPH_REQUIRE(ctx, n >= 0 && n < ((int64_t)1 << 31), "sph_bin_array: bad n %lld",
                (long long)n);
    SPH_REQUIRE(ctx, n == 0 || (x && y && z && perm && sx && sy && sz && cell_of),
                "sph_bin_array: null array");
    int rc = sph_upload_key_tables(ctx, grid);
    if (rc) return rc;
    GridDev g = make_griddev(ctx, grid);
    const int64_t nc = grid->ncells;
    hipLaunchKernelGGL(k_fill_i32, dim3(sph_blocks(nc + 1, NN_THREADS)), dim3(NN_THREADS), 0,
                       ctx->stream, cell_start, nc + 1, 0);
    SPH_LAUNCH_CHECK(ctx, "k_fill_i32");
    if (n == 0) return SPH_OK;
    // the scan below uses the context scratch, so the per-particle rank is
    // parked in sx (n doubles >= n int32) until the scatter has consumed it;
    // k_gather3 then overwrites sx with the sorted x coordinates.
    int32_t *rank = (int32_t *)sx;
    hipLaunchKernelGGL(k_bin_keys, dim3(sph_blocks(n, NN_THREADS)), dim3(NN_THREADS), 0,
                       ctx->stream, g, x, y, z, n, cell_start, cell_of, rank);
    SPH_LAUNCH_CHECK(ctx, "k_bin_keys");
    rc = sph_scan_exclusive_i32(ctx, cell_start, nc);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bin_scatter, dim3(sph_blocks(n, NN_THREADS)), dim3(NN_THREADS), 0,
                       ctx->stream, cell_of, rank, n, cell_start, perm);
    SPH_LAUNCH_CHECK(ctx, "k_bin_scatter");
    hipLaunchKernelGGL(k_cell_sort, dim3(sph_blocks(nc, NN_THREADS)), dim3(NN_THREADS), 0,
                       ctx->stream, cell_start, nc, perm);
    SPH_LAUNCH_CHECK(ctx, "k_cell_sort");
    hipLaunchKernelGGL(k_gather3, dim3(sph_blocks(n, NN_THREADS)), dim3(NN_THREADS), 0,
                       ctx->stream, perm, n, x, y, z, sx, sy, sz);
    SPH_LAUNCH_CHECK(ctx, "k_gather3");
    return SPH_OK;
}

// ---------------------------------------------------------------------------
// neighbour lists
// ---------------------------------------------------------------------------
struct NnpsArgs {
    GridDev g;
    const double *dx, *dy, *dz, *dh;
    int64_t nd_real;
    const int32_t *dperm;  // destination visiting order (may be null)
    int64_t nd_visit;      // length of dperm (or nd_real)
    const int32_t *start;
    const int32_t *sperm;
    const double *sx, *sy, *sz;
    double rs;
    int same;
};

// Visits the candidates of destination i in the reference order and calls
// f(sorted_pos) for each neighbour.
template <typename F>
__device__ __forceinline__ void for_each_neighbour(const NnpsArgs &a, int64_t i, F f)
{
    const GridDev &g = a.g;
    const double px = a.dx[i], py = a.dy[i], pz = a.dz[i];
    const double radius = a.dh[i] * a.rs;
    const double radius2 = radius * radius;
    const int distance = (int)ceil(radius / g.cs);
    const int cx = cell_coord(px, g.cs) - g.cmin0;
    const int cy = cell_coord(py, g.cs) - g.cmin1;
    const int cz = cell_coord(pz, g.cs) - g.cmin2;
    const int x0 = max(cx - distance, 0), x1 = min(cx + distance, g.d0 - 1);
    const int y0 = max(cy - distance, 0), y1 = min(cy + distance, g.d1 - 1);
    const int z0 = max(cz - distance, 0), z1 = min(cz + distance, g.d2 - 1);
    const int32_t *tx = g.tab, *ty = g.tab + g.d0, *tz = g.tab + g.d0 + g.d1;
    for (int X = x0; X <= x1; ++X) {
        const int32_t kx = tx[X];
        for (int Y = y0; Y <= y1; ++Y) {
            const int32_t kxy = kx + ty[Y];
            for (int Z = z0; Z <= z1; ++Z) {
                const int32_t key = kxy + tz[Z];
                const int32_t s = a.start[key], e = a.start[key + 1];
                for (int32_t p = s; p < e; ++p) {
                    const double ex = a.sx[p] - px, ey = a.sy[p] - py, ez = a.sz[p] - pz;
                    const double d2 = ex * ex + ey * ey + ez * ez;
                    if (d2 < radius2) f(p);
                }
            }
        }
    }
}

__global__ void __launch_bounds__(NN_THREADS) k_nnps_count(NnpsArgs a, int64_t *__restrict__ counts)
{
    int64_t t = (int64_t)blockIdx.x * NN_THREADS + threadIdx.x;
    if (t >= a.nd_visit) return;
    int64_t i = a.dperm ? (int64_t)a.dperm[t] : t;
    if (i >= a.nd_real) return;
    int64_t c = 0;
    for_each_neighbour(a, i, [&](int32_t) { ++c; });
    counts[i] = c;
}

__global__ void __launch_bounds__(NN_THREADS)
    k_nnps_fill(NnpsArgs a, const int64_t *__restrict__ offsets, int32_t *__restrict__ indices)
{
    int64_t t = (int64_t)blockIdx.x * NN_THREADS + threadIdx.x;
    if (t >= a.nd_visit) return;
    int64_t i = a.dperm ? (int64_t)a.dperm[t] : t;
    if (i >= a.nd_real) return;
    int32_t *row = indices + offsets[i];
    int64_t c = 0, selfpos = -1;
    const int32_t me = (int32_t)i;
    for_each_neighbour(a, i, [&](int32_t p) {
        const int32_t j = a.sperm[p];
        if (selfpos < 0 && j == me) selfpos = c;
        row[c++] = j;
    });
    // keep the particle itself at the last index (nnps.pyx:612-619)
    if (a.same && selfpos >= 0 && selfpos < c - 1) {
        row[selfpos] = row[c - 1];
        row[c - 1] = me;
    }
}

extern "C" int sph_nnps_build(sph_ctx *ctx, const sph_grid *grid, const double *dx,
                              const double *dy, const double *dz, const double *dh,
                              int64_t nd_real, const int32_t *src_cell_start,
                              const int32_t *src_perm, const double *src_sx,
                              const double *src_sy, const double *src_sz, int64_t ns,
                              double radius_scale, int same_array, int64_t *offsets,
                              int32_t *indices, int64_t cap, int64_t *nnz_out)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, grid && offsets && nnz_out, "sph_nnps_build: null argument");
    SPH_REQUIRE(ctx, nd_real >= 0 && nd_real < ((int64_t)1 << 31), "sph_nnps_build: bad nd");
    SPH_REQUIRE(ctx, ns == 0 || (src_cell_start && src_perm && src_sx && src_sy && src_sz),
                "sph_nnps_build: null source arrays");
    SPH_REQUIRE(ctx, nd_real == 0 || (dx && dy && dz && dh), "sph_nnps_build: null dest");
    int rc = sph_upload_key_tables(ctx, grid);
    if (rc) return rc;
    *nnz_out = 0;
    if (nd_real == 0 || ns == 0) {
        hipLaunchKernelGGL(k_fill_i32, dim3(sph_blocks(2 * (nd_real + 1), NN_THREADS)),
                           dim3(NN_THREADS), 0, ctx->stream, (int32_t *)offsets,
                           2 * (nd_real + 1), 0);
        SPH_LAUNCH_CHECK(ctx, "k_fill_i32");
        return SPH_OK;
    }
    NnpsArgs a;
    a.g = make_griddev(ctx, grid);
    a.dx = dx;
    a.dy = dy;
    a.dz = dz;
    a.dh = dh;
    a.nd_real = nd_real;
    a.dperm = nullptr;
    a.nd_visit = nd_real;
    a.start = src_cell_start;
    a.sperm = src_perm;
    a.sx = src_sx;
    a.sy = src_sy;
    a.sz = src_sz;
    a.rs = radius_scale;
    a.same = same_array;
    hipLaunchKernelGGL(k_nnps_count, dim3(sph_blocks(a.nd_visit, NN_THREADS)), dim3(NN_THREADS), 0,
                       ctx->stream, a, offsets);
    SPH_LAUNCH_CHECK(ctx, "k_nnps_count");
    rc = sph_scan_exclusive_i64(ctx, offsets, nd_real);
    if (rc) return rc;
    rc = sph_ensure_pinned(ctx, sizeof(int64_t));
    if (rc) return rc;
    SPH_HIP_CHECK(ctx, hipMemcpyAsync(ctx->pinned, offsets + nd_real, sizeof(int64_t),
                                      hipMemcpyDeviceToHost, ctx->stream));
    SPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    int64_t nnz = *(int64_t *)ctx->pinned;
    *nnz_out = nnz;
    if (nnz > cap || !indices) {
        sph_set_error(ctx, SPH_ENEEDCAP, "neighbour capacity %lld < %lld", (long long)cap,
                      (long long)nnz);
        return SPH_ENEEDCAP;
    }
    hipLaunchKernelGGL(k_nnps_fill, dim3(sph_blocks(a.nd_visit, NN_THREADS)), dim3(NN_THREADS), 0,
                       ctx->stream, a, (const int64_t *)offsets, indices);
    SPH_LAUNCH_CHECK(ctx, "k_nnps_fill");
    return SPH_OK;
}

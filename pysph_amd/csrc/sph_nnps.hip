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
                              int64_t max_cells, int pad, sph_grid *grid)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, grid && narrays >= 0 && xs && ys && zs && hs && ns, "sph_grid_build: null");
    SPH_REQUIRE(ctx, order == SPH_ORDER_MORTON || order == SPH_ORDER_LINEAR,
                "sph_grid_build: bad order %d", order);
    SPH_REQUIRE(ctx, pad >= 0 && pad <= 1024, "sph_grid_build: bad pad %d", pad);
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
        // pad empty cells on both sides: particles that move a little stay
        // inside the grid, which can then be kept (sph_grid_check)
        grid->cmin[a] = (int32_t)fl - pad;
        int64_t d = (int64_t)fh - (int64_t)fl + 1 + 2 * (int64_t)pad;
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
    d.key_base = 0;
    return d;
}

__global__ void k_fill_i32(int32_t *__restrict__ p, int64_t n, int32_t v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// keeping a grid: do the arrays still fit it?
// ---------------------------------------------------------------------------
// flags |= 1: a particle lies within `margin` cells of the grid's edge (the
// grid should be rebuilt before the next binning); 2: a particle lies outside
// the grid (a binning with this grid clamps it to an edge cell: its lists are
// wrong); 8: a non-finite coordinate.  One atomic per wavefront with a flag.
#define GRID_NEAR_EDGE 1
#define GRID_OUTSIDE 2
#define GRID_NONFINITE 8

__global__ void __launch_bounds__(NN_THREADS)
    k_grid_check(GridDev g, int margin, const double *__restrict__ x,
                 const double *__restrict__ y, const double *__restrict__ z, int64_t n,
                 int32_t *__restrict__ flags)
{
    int32_t f = 0;
    for (int64_t i = (int64_t)blockIdx.x * NN_THREADS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * NN_THREADS) {
        const double a = x[i], b = y[i], c = z[i];
        if (!isfinite(a) || !isfinite(b) || !isfinite(c)) {
            f |= GRID_NONFINITE;
            continue;
        }
        const double ca = floor(a / g.cs) - g.cmin0, cb = floor(b / g.cs) - g.cmin1,
                     cc = floor(c / g.cs) - g.cmin2;
        if (ca < 0.0 || cb < 0.0 || cc < 0.0 || ca >= g.d0 || cb >= g.d1 || cc >= g.d2)
            f |= GRID_OUTSIDE;
        else if (ca < margin || cb < margin || cc < margin || ca >= g.d0 - margin ||
                 cb >= g.d1 - margin || cc >= g.d2 - margin)
            f |= GRID_NEAR_EDGE;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) f |= __shfl_xor(f, d, 64);
    if ((threadIdx.x & 63) == 0 && f) atomicOr(flags, f);
}

extern "C" int sph_grid_check(sph_ctx *ctx, int narrays, const double *const *xs,
                              const double *const *ys, const double *const *zs,
                              const int64_t *ns, const sph_grid *grid, int margin,
                              int32_t *flags)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, grid && flags && narrays >= 0 && (narrays == 0 || (xs && ys && zs && ns)),
                "sph_grid_check: null argument");
    SPH_REQUIRE(ctx, margin >= 0 && grid->cell_size > 0.0, "sph_grid_check: bad margin or grid");
    GridDev g = make_griddev(ctx, grid);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, ctx->stream, flags, 1, 0);
    SPH_LAUNCH_CHECK(ctx, "k_fill_i32");
    for (int a = 0; a < narrays; ++a) {
        if (ns[a] <= 0) continue;
        SPH_REQUIRE(ctx, xs[a] && ys[a] && zs[a], "sph_grid_check: null array %d", a);
        int64_t nb = (ns[a] + NN_THREADS - 1) / NN_THREADS;
        if (nb > 1024) nb = 1024;
        hipLaunchKernelGGL(k_grid_check, dim3((unsigned)nb), dim3(NN_THREADS), 0, ctx->stream, g,
                           margin, xs[a], ys[a], zs[a], ns[a], flags);
        SPH_LAUNCH_CHECK(ctx, "k_grid_check");
    }
    return SPH_OK;
}


__global__ void __launch_bounds__(NN_THREADS)
    k_bin_keys(GridDev g, const double *__restrict__ x, const double *__restrict__ y,
               const double *__restrict__ z, int64_t n, int32_t *__restrict__ count,
               int32_t *__restrict__ cell_of, int32_t *__restrict__ rank)
{
    const int64_t i = (int64_t)blockIdx.x * NN_THREADS + threadIdx.x;
    const bool act = i < n;  // the active lanes are a prefix of the wavefront
    int32_t key = -1;
    if (act) {
        int cx = cell_coord(x[i], g.cs) - g.cmin0;
        int cy = cell_coord(y[i], g.cs) - g.cmin1;
        int cz = cell_coord(z[i], g.cs) - g.cmin2;
        // every particle lies in the grid by construction (sph_grid_build);
        // clamp defensively so a stale grid can never index outside the tables
        cx = min(max(cx, 0), g.d0 - 1);
        cy = min(max(cy, 0), g.d1 - 1);
        cz = min(max(cz, 0), g.d2 - 1);
        key = cell_key(g, cx, cy, cz);
        cell_of[i] = key;
    }
    // consecutive lanes with the same cell (neighbours in memory are mostly
    // neighbours in space) take their slots with one atomic per run; slots
    // only need to be distinct within a cell (k_bin_rank orders them)
    const int lane = __lane_id();
    const int32_t prev = __shfl_up(key, 1, 64);
    const bool head = act && (lane == 0 || prev != key);
    const uint64_t heads = __builtin_amdgcn_ballot_w64(head);
    const uint64_t acts = __builtin_amdgcn_ballot_w64(act);
    const uint64_t le = heads & ((2ull << lane) - 1ull);  // heads at or below my lane
    const int start = le ? 63 - __clzll(le) : 0;
    const uint64_t above = heads & ~((2ull << start) - 1ull);
    const int next = above ? __ffsll((long long)above) - 1 : __popcll(acts);
    int32_t base = 0;
    if (head) base = atomicAdd(&count[key], next - start);
    base = __shfl(base, start, 64);
    if (act) rank[i] = base + (lane - start);
}

__global__ void __launch_bounds__(NN_THREADS)
    k_bin_scatter(const int32_t *__restrict__ cell_of, const int32_t *__restrict__ slot,
                  int64_t n, const int32_t *__restrict__ start, int32_t *__restrict__ tmp,
                  int32_t *__restrict__ nlarge)
{
    int64_t i = (int64_t)blockIdx.x * NN_THREADS + threadIdx.x;
    if (i == 0) *nlarge = 0;  // the large-cell counter of k_bin_rank, which runs next
    if (i >= n) return;
    tmp[start[cell_of[i]] + slot[i]] = (int32_t)i;
}

// Cells with at most RANK_DIRECT members are ordered by counting (each
// member scans its cell); larger ones by k_bin_rank_large, a merge sort of
// the cell's members, so a cell holding most of the particles (a large h
// sets the cell size, or a clustered input) costs O(m log^2 m), not O(m^2).
#define RANK_DIRECT 1024
#define RANK_CHUNK 2048        // members sorted in LDS per pass of k_bin_rank_large
#define RANK_LARGE_GRID 256    // blocks of k_bin_rank_large (one large cell at a time each)

// Ascending particle index inside every cell (the atomics above hand out
// slots in arbitrary order): particle i's position is the number of members
// of its cell with a smaller index.  Writes perm and its inverse rank.  The
// members of a cell with more than RANK_DIRECT of them are left to
// k_bin_rank_large: the member holding slot 0 lists the cell there.
__global__ void __launch_bounds__(NN_THREADS)
    k_bin_rank(const int32_t *__restrict__ cell_of, int64_t n, const int32_t *__restrict__ start,
               const int32_t *__restrict__ tmp, int32_t trash, int32_t *__restrict__ perm,
               int32_t *__restrict__ rank, int32_t *__restrict__ large, int32_t *__restrict__ nlarge)
{
    int64_t i = (int64_t)blockIdx.x * NN_THREADS + threadIdx.x;
    if (i >= n) return;
    const int32_t key = cell_of[i];
    const int32_t s = start[key], e = start[key + 1];
    int32_t r = 0;
    if (key == trash) {
        r = rank[i];  // unordered: the atomic slot (no lists are built for this cell)
    } else if (e - s > RANK_DIRECT) {
        if (rank[i] == 0) large[atomicAdd(nlarge, 1)] = key;  // slot 0: once per cell
        return;
    } else {
        // eight members' loads in flight before the first compare (a loop of
        // one load at a time waits out the load latency per member)
        int32_t a = s;
        for (; a + 8 <= e; a += 8) {
            int32_t v[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) v[k] = tmp[a + k];
#pragma unroll
            for (int k = 0; k < 8; ++k) r += v[k] < (int32_t)i;
        }
        for (; a < e; ++a) r += tmp[a] < (int32_t)i;
    }
    perm[s + r] = (int32_t)i;
    rank[i] = s + r;
}

// number of the n ascending values of a that are smaller than v
__device__ __forceinline__ int64_t lower_bound_i32(const int32_t *a, int64_t n, int32_t v)
{
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// The members of each listed large cell in ascending order: chunks of
// RANK_CHUNK sorted in LDS (bitonic), then pairwise merges of runs of width
// w, 2w, ... between tmp[s..e) and perm[s..e) (a member's place in the
// merged run is its place in its own run plus the number of smaller members
// of the other run; members are distinct indices).  The result ends in
// perm[s..e) and rank is its inverse.  One block per cell at a time.
__global__ void __launch_bounds__(NN_THREADS)
    k_bin_rank_large(const int32_t *__restrict__ start, int32_t *__restrict__ tmp,
                     const int32_t *__restrict__ large, const int32_t *__restrict__ nlarge,
                     int32_t *__restrict__ perm, int32_t *__restrict__ rank)
{
    __shared__ int32_t buf[RANK_CHUNK];
    const int32_t nl = *nlarge;
    const int tid = threadIdx.x;
    for (int32_t c = blockIdx.x; c < nl; c += gridDim.x) {
        const int32_t key = large[c];
        const int32_t s = start[key];
        const int64_t m = (int64_t)start[key + 1] - s;
        int32_t *a = tmp + s;
        int32_t *b = perm + s;
        for (int64_t c0 = 0; c0 < m; c0 += RANK_CHUNK) {
            const int64_t len = m - c0 < RANK_CHUNK ? m - c0 : RANK_CHUNK;
            for (int k = tid; k < RANK_CHUNK; k += NN_THREADS)
                buf[k] = k < len ? a[c0 + k] : INT32_MAX;
            __syncthreads();
            for (int k = 2; k <= RANK_CHUNK; k <<= 1) {
                for (int j = k >> 1; j > 0; j >>= 1) {
                    for (int t = tid; t < RANK_CHUNK; t += NN_THREADS) {
                        const int u = t ^ j;
                        if (u > t) {
                            const int32_t x = buf[t], y = buf[u];
                            if ((x > y) == ((t & k) == 0)) {
                                buf[t] = y;
                                buf[u] = x;
                            }
                        }
                    }
                    __syncthreads();
                }
            }
            for (int k = tid; k < len; k += NN_THREADS) a[c0 + k] = buf[k];
            __syncthreads();
        }
        int32_t *src = a, *dst = b;
        for (int64_t w = RANK_CHUNK; w < m; w *= 2) {
            for (int64_t e = tid; e < m; e += NN_THREADS) {
                const int64_t p0 = e / (2 * w) * (2 * w);
                const int64_t off = e - p0;
                const int64_t lenA = m - p0 < w ? m - p0 : w;
                const int64_t restB = m - p0 - w;
                const int64_t lenB = restB <= 0 ? 0 : (restB < w ? restB : w);
                const int32_t v = src[e];
                const int64_t pos = off < w ? off + lower_bound_i32(src + p0 + w, lenB, v)
                                            : (off - w) + lower_bound_i32(src + p0, lenA, v);
                dst[p0 + pos] = v;
            }
            __syncthreads();
            int32_t *t = src;
            src = dst;
            dst = t;
        }
        for (int64_t e = tid; e < m; e += NN_THREADS) {
            const int32_t v = src[e];
            if (src != b) b[e] = v;
            rank[v] = s + (int32_t)e;
        }
        __syncthreads();
    }
}

// Cell starts and the list of non-empty cells from the cell counts, in two
// launches: per-tile totals (members, non-empty cells), then each tile adds
// the totals of the tiles before it to its own scan and writes the starts,
// its non-empty cells in key order and, in the last tile, the totals.
// Out of place (counts are not overwritten); up to BSCAN_MAX_TILES tiles,
// beyond that the multi-level scans run instead.
#define BSCAN_THREADS 256
#define BSCAN_ITEMS 16
#define BSCAN_TILE (BSCAN_THREADS * BSCAN_ITEMS)
#define BSCAN_MAX_TILES 2048

__global__ void __launch_bounds__(NN_THREADS)
    k_copy_i32(const int32_t *__restrict__ src, int64_t n, int32_t *__restrict__ dst)
{
    const int64_t i = (int64_t)blockIdx.x * NN_THREADS + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// (members, non-empty cells) of a block's BSCAN_TILE cells, summed over the block
__device__ __forceinline__ int2 bscan_block_sum(int32_t a, int32_t b, int2 *s_part)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        a += __shfl_xor(a, d, 64);
        b += __shfl_xor(b, d, 64);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) s_part[wave] = make_int2(a, b);
    __syncthreads();
    int2 t = make_int2(0, 0);
    for (int q = 0; q < BSCAN_THREADS / 64; ++q) {
        t.x += s_part[q].x;
        t.y += s_part[q].y;
    }
    __syncthreads();
    return t;
}

__global__ void __launch_bounds__(BSCAN_THREADS)
    k_bin_tile_sums(const int32_t *__restrict__ counts, int64_t nc, int2 *__restrict__ tile_sum)
{
    __shared__ int2 s_part[BSCAN_THREADS / 64];
    const int64_t base = (int64_t)blockIdx.x * BSCAN_TILE + (int64_t)threadIdx.x * BSCAN_ITEMS;
    int32_t a = 0, b = 0;
#pragma unroll
    for (int k = 0; k < BSCAN_ITEMS; ++k) {
        const int32_t c = base + k < nc ? counts[base + k] : 0;
        a += c;
        b += c > 0;
    }
    const int2 t = bscan_block_sum(a, b, s_part);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = t;
}

__global__ void __launch_bounds__(BSCAN_THREADS)
    k_bin_tile_scan(const int32_t *__restrict__ counts, int64_t nc,
                    const int2 *__restrict__ tile_sum, int32_t *__restrict__ start,
                    int32_t *__restrict__ cells, int32_t *__restrict__ ncells_out,
                    int32_t *__restrict__ ncells_copy)
{
    __shared__ int2 s_part[BSCAN_THREADS / 64];
    __shared__ int2 s_wave[BSCAN_THREADS / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // totals of the tiles before this one, eight loads in flight per thread
    int32_t pa = 0, pb = 0;
    for (int64_t q0 = 0; q0 < (int64_t)blockIdx.x; q0 += 8 * BSCAN_THREADS) {
        int2 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int64_t q = q0 + (int64_t)u * BSCAN_THREADS + tid;
            v[u] = q < (int64_t)blockIdx.x ? tile_sum[q] : make_int2(0, 0);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            pa += v[u].x;
            pb += v[u].y;
        }
    }
    const int2 pre = bscan_block_sum(pa, pb, s_part);
    const int64_t base = (int64_t)blockIdx.x * BSCAN_TILE + (int64_t)tid * BSCAN_ITEMS;
    int32_t c[BSCAN_ITEMS];
    int32_t a = 0, b = 0;
#pragma unroll
    for (int k = 0; k < BSCAN_ITEMS; ++k) {
        c[k] = base + k < nc ? counts[base + k] : 0;
        a += c[k];
        b += c[k] > 0;
    }
    // exclusive prefix of this thread inside the tile
    int32_t ia = a, ib = b;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int32_t ua = __shfl_up(ia, d, 64), ub = __shfl_up(ib, d, 64);
        if (lane >= d) {
            ia += ua;
            ib += ub;
        }
    }
    if (lane == 63) s_wave[wave] = make_int2(ia, ib);
    __syncthreads();
    int32_t ea = pre.x + ia - a, eb = pre.y + ib - b;
    int32_t ta = 0, tb = 0;  // the tile's totals
    for (int q = 0; q < BSCAN_THREADS / 64; ++q) {
        const int2 v = s_wave[q];
        if (q < wave) {
            ea += v.x;
            eb += v.y;
        }
        ta += v.x;
        tb += v.y;
    }
#pragma unroll
    for (int k = 0; k < BSCAN_ITEMS; ++k) {
        const int64_t key = base + k;
        if (key < nc) {
            start[key] = ea;
            if (c[k] > 0) cells[eb++] = (int32_t)key;
        }
        ea += c[k];
    }
    if (blockIdx.x == gridDim.x - 1 && tid == 0) {
        start[nc] = pre.x + ta;
        *ncells_out = pre.y + tb;
        if (ncells_copy) *ncells_copy = pre.y + tb;
    }
}

__global__ void __launch_bounds__(NN_THREADS)
    k_cell_flags(const int32_t *__restrict__ start, int64_t ncells, int32_t *__restrict__ nonempty)
{
    int64_t c = (int64_t)blockIdx.x * NN_THREADS + threadIdx.x;
    if (c >= ncells) return;
    nonempty[c] = start[c + 1] > start[c] ? 1 : 0;
}

__global__ void __launch_bounds__(NN_THREADS)
    k_compact_cells(const int32_t *__restrict__ start, const int32_t *__restrict__ slot,
                    int64_t ncells, int32_t *__restrict__ cells)
{
    int64_t c = (int64_t)blockIdx.x * NN_THREADS + threadIdx.x;
    if (c >= ncells) return;
    if (start[c + 1] > start[c]) cells[slot[c]] = (int32_t)c;
}

static int read_i64(sph_ctx *ctx, const void *dev, int64_t *out, bool is32)
{
    int rc = sph_ensure_pinned(ctx, sizeof(int64_t));
    if (rc) return rc;
    SPH_HIP_CHECK(ctx, hipMemcpyAsync(ctx->pinned, dev, is32 ? sizeof(int32_t) : sizeof(int64_t),
                                      hipMemcpyDeviceToHost, ctx->stream));
    SPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    *out = is32 ? (int64_t) * (int32_t *)ctx->pinned : *(int64_t *)ctx->pinned;
    return SPH_OK;
}

static size_t bin_large_off(int64_t n, int64_t nc, size_t scratch_off)
{
    size_t part = sph_scan_scratch_bytes(nc > n ? nc : n, sizeof(int32_t));
    const size_t tiles = sizeof(int2) * (size_t)(nc / BSCAN_TILE + 1) + 256;
    if (tiles > part) part = tiles;
    return (scratch_off + part + 255) & ~(size_t)255;
}

size_t sph_bin_sort_scratch_end(int64_t n, int64_t nc, size_t scratch_off)
{
    return bin_large_off(n, nc, scratch_off) +
           sizeof(int32_t) * (size_t)(n / (RANK_DIRECT + 1) + 2);
}

// slot[0..nc) holds the cell counts on entry (k_bin_keys); cell_start[0..nc]
// receives the starts, cells the non-empty cell keys ascending and slot[nc]
// their number
int sph_bin_sort(sph_ctx *ctx, int64_t n, int64_t nc, const int32_t *cell_of, int32_t *rank,
                 int32_t *cell_start, int32_t *perm, int32_t *cells, int32_t *slot, int32_t *tmp,
                 int32_t trash, size_t scratch_off, int32_t *ncells_copy)
{
    // the large-cell list (a counter and at most n / (RANK_DIRECT + 1) keys)
    // after the scan partials; the callers size the scratch for it first
    // (sph_bin_sort_scratch_end), so this never moves it under their pointers
    const size_t large_off = bin_large_off(n, nc, scratch_off);
    int rc = sph_ensure_scratch(ctx, sph_bin_sort_scratch_end(n, nc, scratch_off));
    if (rc) return rc;
    int32_t *nlarge = (int32_t *)((char *)ctx->scratch + large_off);
    int32_t *large = nlarge + 1;
    const int64_t ntiles = (nc + BSCAN_TILE - 1) / BSCAN_TILE;
    if (ntiles <= BSCAN_MAX_TILES) {
        // starts, the compact list of non-empty cell keys and its count
        int2 *tile_sum = (int2 *)((char *)ctx->scratch + ((scratch_off + 255) & ~(size_t)255));
        hipLaunchKernelGGL(k_bin_tile_sums, dim3((unsigned)ntiles), dim3(BSCAN_THREADS), 0,
                           ctx->stream, slot, nc, tile_sum);
        SPH_LAUNCH_CHECK(ctx, "k_bin_tile_sums");
        hipLaunchKernelGGL(k_bin_tile_scan, dim3((unsigned)ntiles), dim3(BSCAN_THREADS), 0,
                           ctx->stream, slot, nc, (const int2 *)tile_sum, cell_start, cells,
                           slot + nc, ncells_copy);
        SPH_LAUNCH_CHECK(ctx, "k_bin_tile_scan");
    } else {
        hipLaunchKernelGGL(k_copy_i32, dim3(sph_blocks(nc, NN_THREADS)), dim3(NN_THREADS), 0,
                           ctx->stream, slot, nc, cell_start);
        SPH_LAUNCH_CHECK(ctx, "k_copy_i32");
        rc = sph_scan_exclusive_i32_at(ctx, cell_start, nc, scratch_off);
        if (rc) return rc;
        // non-empty flags -> slots (scan) -> compact list of cell keys
        hipLaunchKernelGGL(k_cell_flags, dim3(sph_blocks(nc, NN_THREADS)), dim3(NN_THREADS), 0,
                           ctx->stream, cell_start, nc, slot);
        SPH_LAUNCH_CHECK(ctx, "k_cell_flags");
        rc = sph_scan_exclusive_i32_at(ctx, slot, nc, scratch_off);
        if (rc) return rc;
        hipLaunchKernelGGL(k_compact_cells, dim3(sph_blocks(nc, NN_THREADS)), dim3(NN_THREADS),
                           0, ctx->stream, cell_start, slot, nc, cells);
        SPH_LAUNCH_CHECK(ctx, "k_compact_cells");
        if (ncells_copy)
            SPH_HIP_CHECK(ctx, hipMemcpyAsync(ncells_copy, slot + nc, sizeof(int32_t),
                                              hipMemcpyDeviceToDevice, ctx->stream));
    }
    hipLaunchKernelGGL(k_bin_scatter, dim3(sph_blocks(n > 0 ? n : 1, NN_THREADS)),
                       dim3(NN_THREADS), 0, ctx->stream, cell_of, rank, n, cell_start, tmp,
                       nlarge);
    SPH_LAUNCH_CHECK(ctx, "k_bin_scatter");
    hipLaunchKernelGGL(k_bin_rank, dim3(sph_blocks(n, NN_THREADS)), dim3(NN_THREADS), 0,
                       ctx->stream, cell_of, n, cell_start, tmp, trash, perm, rank, large,
                       nlarge);
    SPH_LAUNCH_CHECK(ctx, "k_bin_rank");
    hipLaunchKernelGGL(k_bin_rank_large, dim3(RANK_LARGE_GRID), dim3(NN_THREADS), 0,
                       ctx->stream, cell_start, tmp, large, nlarge, perm, rank);
    SPH_LAUNCH_CHECK(ctx, "k_bin_rank_large");
    return SPH_OK;
}

// ncells_out (host, synchronising) or ncells_dev (device, stream-ordered)
static int bin_array(sph_ctx *ctx, const sph_grid *grid, const double *x, const double *y,
                     const double *z, int64_t n, int32_t *cell_start, int32_t *perm,
                     int32_t *rank, int32_t *cell_of, int32_t *cells, int64_t *ncells_out,
                     int32_t *ncells_dev)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, grid && grid->ncells > 0 && cell_start && (ncells_out || ncells_dev),
                "sph_bin_array: bad grid / null output");
    SPH_REQUIRE(ctx, n >= 0 && n < ((int64_t)1 << 31), "sph_bin_array: bad n %lld",
                (long long)n);
    SPH_REQUIRE(ctx, n == 0 || (x && y && z && perm && rank && cell_of && cells),
                "sph_bin_array: null array");
    int rc = sph_upload_key_tables(ctx, grid);
    if (rc) return rc;
    GridDev g = make_griddev(ctx, grid);
    const int64_t nc = grid->ncells;
    if (ncells_out) *ncells_out = 0;
    if (n == 0) {
        hipLaunchKernelGGL(k_fill_i32, dim3(sph_blocks(nc + 1, NN_THREADS)), dim3(NN_THREADS),
                           0, ctx->stream, cell_start, nc + 1, 0);
        SPH_LAUNCH_CHECK(ctx, "k_fill_i32");
        if (ncells_dev) {
            hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, ctx->stream, ncells_dev, 1, 0);
            SPH_LAUNCH_CHECK(ctx, "k_fill_i32");
        }
        return SPH_OK;
    }
    // scratch: [non-empty flags / slots (nc+1) | unsorted cell members (n) | scan partials |
    // cells with many members];
    // the scan of cell_start runs before either region is written
    const size_t slot_bytes = (sizeof(int32_t) * (size_t)(nc + 1) + 511) & ~(size_t)255;
    const size_t tmp_bytes = (sizeof(int32_t) * (size_t)n + 255) & ~(size_t)255;
    rc = sph_ensure_scratch(ctx, sph_bin_sort_scratch_end(n, nc, slot_bytes + tmp_bytes));
    if (rc) return rc;
    int32_t *slot = (int32_t *)ctx->scratch;
    int32_t *tmp = (int32_t *)((char *)ctx->scratch + slot_bytes);
    // cell counts into slot (zeroed first); rank first holds the slot inside
    // the cell handed out by the atomics, sph_bin_sort overwrites it with the
    // sorted position
    hipLaunchKernelGGL(k_fill_i32, dim3(sph_blocks(nc + 1, NN_THREADS)), dim3(NN_THREADS), 0,
                       ctx->stream, slot, nc + 1, 0);
    SPH_LAUNCH_CHECK(ctx, "k_fill_i32");
    hipLaunchKernelGGL(k_bin_keys, dim3(sph_blocks(n, NN_THREADS)), dim3(NN_THREADS), 0,
                       ctx->stream, g, x, y, z, n, slot, cell_of, rank);
    SPH_LAUNCH_CHECK(ctx, "k_bin_keys");
    // a device-side binning's count also goes to ncells_dev
    rc = sph_bin_sort(ctx, n, nc, cell_of, rank, cell_start, perm, cells, slot, tmp, -1,
                      slot_bytes + tmp_bytes, ncells_dev);
    if (rc) return rc;
    if (ncells_dev) return SPH_OK;
    return read_i64(ctx, slot + nc, ncells_out, true);
}

extern "C" int sph_bin_array(sph_ctx *ctx, const sph_grid *grid, const double *x,
                             const double *y, const double *z, int64_t n, int32_t *cell_start,
                             int32_t *perm, int32_t *rank, int32_t *cell_of, int32_t *cells,
                             int64_t *ncells_out)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, ncells_out, "sph_bin_array: null ncells_out");
    return bin_array(ctx, grid, x, y, z, n, cell_start, perm, rank, cell_of, cells, ncells_out,
                     nullptr);
}

extern "C" int sph_bin_array_dev(sph_ctx *ctx, const sph_grid *grid, const double *x,
                                 const double *y, const double *z, int64_t n,
                                 int32_t *cell_start, int32_t *perm, int32_t *rank,
                                 int32_t *cell_of, int32_t *cells, int32_t *ncells_dev)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, ncells_dev, "sph_bin_array_dev: null ncells_dev");
    return bin_array(ctx, grid, x, y, z, n, cell_start, perm, rank, cell_of, cells, nullptr,
                     ncells_dev);
}

__global__ void __launch_bounds__(NN_THREADS)
    k_gather_sorted(int64_t n, const int32_t *__restrict__ perm, sph_prop_list pl)
{
    int64_t t = (int64_t)blockIdx.x * NN_THREADS + threadIdx.x;
    if (t >= n) return;
    const int32_t i = perm[t];
    const int64_t o = t * pl.dst_stride;
    double v[SPH_MAX_PROPS];  // all loads before the first store (in-order retirement)
#pragma unroll
    for (int k = 0; k < SPH_MAX_PROPS; ++k) v[k] = k < pl.nprops ? pl.src[k][i] : 0.0;
#pragma unroll
    for (int k = 0; k < SPH_MAX_PROPS; ++k)
        if (k < pl.nprops) pl.dst[k][o] = v[k];
}

__global__ void __launch_bounds__(NN_THREADS)
    k_gather_sorted_f32(int64_t n, const int32_t *__restrict__ perm, sph_prop_list_f32 pl)
{
    int64_t t = (int64_t)blockIdx.x * NN_THREADS + threadIdx.x;
    if (t >= n) return;
    const int32_t i = perm[t];
    const int64_t o = t * pl.dst_stride;
    double v[SPH_MAX_PROPS];  // all loads before the first store (in-order retirement)
#pragma unroll
    for (int k = 0; k < SPH_MAX_PROPS; ++k) v[k] = k < pl.nprops ? pl.src[k][i] : 0.0;
#pragma unroll
    for (int k = 0; k < SPH_MAX_PROPS; ++k)
        if (k < pl.nprops) pl.dst[k][o] = (float)v[k];
}

// the same records written from the original order: thread i reads field k
// of particle i (coalesced) and writes it into row rank[i].  A row is 96
// contiguous bytes, so the scattered stores of one particle fill its row in
// L2, while the gather above reads each field with one cache line per lane.
// one particle's fields v into row o of every destination of pl
__device__ __forceinline__ void scatter_row(const sph_prop_list &pl, const double *v, int64_t o)
{
    bool taken = false;  // field k went out with field k - 1
#pragma unroll
    for (int k = 0; k < SPH_MAX_PROPS; ++k) {
        if (k >= pl.nprops) break;
        if (taken) {
            taken = false;
            continue;
        }
        double *d = pl.dst[k] + o;
        // two fields that are adjacent 16-byte-aligned words of the row go
        // out as one 16-byte store (the test is the same for every lane)
        if (k + 1 < pl.nprops && pl.dst[k + 1] == pl.dst[k] + 1 &&
            ((uintptr_t)d & 15) == 0) {
            *reinterpret_cast<double2 *>(d) = make_double2(v[k], v[k + 1]);
            taken = true;
        } else {
            *d = v[k];
        }
    }
}

__global__ void __launch_bounds__(NN_THREADS)
    k_scatter_sorted(int64_t n, const int32_t *__restrict__ rank, sph_prop_list pl)
{
    int64_t i = (int64_t)blockIdx.x * NN_THREADS + threadIdx.x;
    if (i >= n) return;
    const int64_t o = (int64_t)rank[i] * pl.dst_stride;
    // every field is loaded before the first store: memory operations retire
    // in order, so a load issued after a store would wait for that store
    double v[SPH_MAX_PROPS];
#pragma unroll
    for (int k = 0; k < SPH_MAX_PROPS; ++k) v[k] = k < pl.nprops ? pl.src[k][i] : 0.0;
    scatter_row(pl, v, o);
}

// two prop lists of their own row strides in one pass (the sorted records
// and the compact rows from the same original-order fields): every load of
// both lists before the first store; a field both lists read is loaded twice
// (the second from cache)
__global__ void __launch_bounds__(NN_THREADS)
    k_scatter_sorted2(int64_t n, const int32_t *__restrict__ rank, sph_prop_list pa,
                      sph_prop_list pb)
{
    int64_t i = (int64_t)blockIdx.x * NN_THREADS + threadIdx.x;
    if (i >= n) return;
    const int64_t r = rank[i];
    double va[SPH_MAX_PROPS], vb[SPH_MAX_PROPS];
#pragma unroll
    for (int k = 0; k < SPH_MAX_PROPS; ++k) va[k] = k < pa.nprops ? pa.src[k][i] : 0.0;
#pragma unroll
    for (int k = 0; k < SPH_MAX_PROPS; ++k) vb[k] = k < pb.nprops ? pb.src[k][i] : 0.0;
    scatter_row(pa, va, r * pa.dst_stride);
    scatter_row(pb, vb, r * pb.dst_stride);
}

__global__ void __launch_bounds__(NN_THREADS)
    k_scatter_sorted_f32(int64_t n, const int32_t *__restrict__ rank, sph_prop_list_f32 pl)
{
    int64_t i = (int64_t)blockIdx.x * NN_THREADS + threadIdx.x;
    if (i >= n) return;
    const int64_t o = (int64_t)rank[i] * pl.dst_stride;
    double v[SPH_MAX_PROPS];  // all loads before the first store (in-order retirement)
#pragma unroll
    for (int k = 0; k < SPH_MAX_PROPS; ++k) v[k] = k < pl.nprops ? pl.src[k][i] : 0.0;
#pragma unroll
    for (int k = 0; k < SPH_MAX_PROPS; ++k)
        if (k < pl.nprops) pl.dst[k][o] = (float)v[k];
}

template <typename PL, typename K>
static int scatter_sorted(sph_ctx *ctx, const char *name, K kern, int64_t n, const int32_t *rank,
                          const PL *props)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, props && props->nprops >= 0 && props->nprops <= SPH_MAX_PROPS &&
                         props->dst_stride >= 1,
                "%s: bad props", name);
    if (n <= 0 || props->nprops == 0) return SPH_OK;
    SPH_REQUIRE(ctx, rank, "%s: null rank", name);
    for (int k = 0; k < props->nprops; ++k)
        SPH_REQUIRE(ctx, props->src[k] && props->dst[k], "%s: null prop %d", name, k);
    hipLaunchKernelGGL(kern, dim3(sph_blocks(n, NN_THREADS)), dim3(NN_THREADS), 0, ctx->stream,
                       n, rank, *props);
    SPH_LAUNCH_CHECK(ctx, name);
    return SPH_OK;
}

extern "C" int sph_scatter_sorted(sph_ctx *ctx, int64_t n, const int32_t *rank,
                                  const sph_prop_list *props)
{
    return scatter_sorted(ctx, "sph_scatter_sorted", k_scatter_sorted, n, rank, props);
}

extern "C" int sph_scatter_sorted2(sph_ctx *ctx, int64_t n, const int32_t *rank,
                                   const sph_prop_list *props, const sph_prop_list *props2)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, props2 && props2->nprops >= 0 && props2->nprops <= SPH_MAX_PROPS &&
                         props2->dst_stride >= 1,
                "sph_scatter_sorted2: bad second props");
    if (props2->nprops == 0)
        return scatter_sorted(ctx, "sph_scatter_sorted2", k_scatter_sorted, n, rank, props);
    SPH_REQUIRE(ctx, props && props->nprops >= 0 && props->nprops <= SPH_MAX_PROPS &&
                         props->dst_stride >= 1,
                "sph_scatter_sorted2: bad props");
    if (n <= 0) return SPH_OK;
    SPH_REQUIRE(ctx, rank, "sph_scatter_sorted2: null rank");
    for (int k = 0; k < props->nprops; ++k)
        SPH_REQUIRE(ctx, props->src[k] && props->dst[k], "sph_scatter_sorted2: null prop %d", k);
    for (int k = 0; k < props2->nprops; ++k)
        SPH_REQUIRE(ctx, props2->src[k] && props2->dst[k],
                    "sph_scatter_sorted2: null second prop %d", k);
    hipLaunchKernelGGL(k_scatter_sorted2, dim3(sph_blocks(n, NN_THREADS)), dim3(NN_THREADS), 0,
                       ctx->stream, n, rank, *props, *props2);
    SPH_LAUNCH_CHECK(ctx, "sph_scatter_sorted2");
    return SPH_OK;
}

extern "C" int sph_scatter_sorted_f32(sph_ctx *ctx, int64_t n, const int32_t *rank,
                                      const sph_prop_list_f32 *props)
{
    return scatter_sorted(ctx, "sph_scatter_sorted_f32", k_scatter_sorted_f32, n, rank, props);
}

extern "C" int sph_gather_sorted_f32(sph_ctx *ctx, int64_t n, const int32_t *perm,
                                     const sph_prop_list_f32 *props)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, props && props->nprops >= 0 && props->nprops <= SPH_MAX_PROPS &&
                         props->dst_stride >= 1,
                "sph_gather_sorted_f32: bad props");
    if (n <= 0 || props->nprops == 0) return SPH_OK;
    SPH_REQUIRE(ctx, perm, "sph_gather_sorted_f32: null perm");
    for (int k = 0; k < props->nprops; ++k)
        SPH_REQUIRE(ctx, props->src[k] && props->dst[k], "sph_gather_sorted_f32: null prop %d",
                    k);
    hipLaunchKernelGGL(k_gather_sorted_f32, dim3(sph_blocks(n, NN_THREADS)), dim3(NN_THREADS), 0,
                       ctx->stream, n, perm, *props);
    SPH_LAUNCH_CHECK(ctx, "k_gather_sorted_f32");
    return SPH_OK;
}

extern "C" int sph_gather_sorted(sph_ctx *ctx, int64_t n, const int32_t *perm,
                                 const sph_prop_list *props)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, props && props->nprops >= 0 && props->nprops <= SPH_MAX_PROPS &&
                         props->dst_stride >= 1,
                "sph_gather_sorted: bad props");
    if (n <= 0 || props->nprops == 0) return SPH_OK;
    SPH_REQUIRE(ctx, perm, "sph_gather_sorted: null perm");
    for (int k = 0; k < props->nprops; ++k)
        SPH_REQUIRE(ctx, props->src[k] && props->dst[k], "sph_gather_sorted: null prop %d", k);
    hipLaunchKernelGGL(k_gather_sorted, dim3(sph_blocks(n, NN_THREADS)), dim3(NN_THREADS), 0,
                       ctx->stream, n, perm, *props);
    SPH_LAUNCH_CHECK(ctx, "k_gather_sorted");
    return SPH_OK;
}

// ---------------------------------------------------------------------------
// neighbour lists (sorted space): one 64-lane wavefront per destination cell
// ---------------------------------------------------------------------------
#define CW 64          // lanes per wavefront = threads per block of the cell kernel
#define MAXSEG 64      // stencil cells handled per batch (27 for d = 1)

#define NN_MAXCH 16    // candidate chunks (of 64) whose hit masks the count pass keeps
#define NN_CELL_GRID (1 << 20)  // blocks of the count pass (one cell at a time each)
#define NN_OVF_GRID 2048    // blocks of the per-cell fill of flagged cells

struct CellArgs {
    GridDev g;
    // records of the destination and source arrays: double, or float when
    // f32 (then the destination cell coordinates come from ox, oy, oz, the
    // double coordinates the binning used, in original order)
    int f32;
    const float *drec32, *srec32;
    const double *ox, *oy, *oz;
    const double *drec;
    const int32_t *dperm, *dstart, *dcells;
    int64_t ncell_list, nd_real;
    const int32_t *sstart;
    const double *srec;
    double rs;
    int same;
    // hit masks of the count pass: chunk c of sorted destination t at
    // masks[c * mstride + t]; cell_ovf[w] = the chunks of cell w's stencil,
    // or -1 when it has more than maxch chunks or more than one stencil batch
    // (its fill recomputes the distance tests)
    uint64_t *masks;
    int64_t mstride;
    int32_t *cell_ovf;
    // per cell: the sorted source of each stencil candidate q (q <
    // NN_MAXCH*64, cells without the overflow flag); per destination: its cell
    int32_t *cand;
    int32_t *dcell_of;
    int maxch;  // chunks kept per destination (<= NN_MAXCH); cand holds 64*maxch per cell
    // batched simulations: per-simulation grids (grids[sim_of[original
    // index]] replaces g), the number of destination cells read on the
    // device, and the capacity the fill checks on the device (lists that do
    // not fit set bit 0 of *status and are not written)
    const GridDev *grids;
    const int32_t *sim_of;
    const int32_t *key_sim;  // or the simulation of each cell key (NULL: via sim_of)
    const int32_t *ncell_dev;
    int64_t nslices;
    int64_t cap;
    int32_t *status;
    int32_t trash;  // cell key that gets no lists (particles of void simulations), or -1
    // cells flagged by the count pass, for the per-cell fill
    int32_t *ovf_list;
    int32_t *novf;
    // destination cells per block of the count pass (its grid is the cell
    // count over this, at most NN_CELL_GRID blocks)
    int cells_per_block;
    // device-side plain builds: the fill also writes totals = {slots, nnz,
    // status} (and sets the status bits) from its first thread
    int64_t *totals;
    const int64_t *nnz_dev;
    int64_t cell_bound;
    // symmetric variable-h lists: source permutation and real count (the
    // appended part takes real sources only, by original index), the
    // stencil reach of the largest radius, the own-query entries of every
    // destination and the appended part's hit masks (same layout as masks)
    const int32_t *sperm;
    int64_t ns_real;
    int reach;
    int32_t *fcount;
    uint64_t *amasks;
    int32_t *candj;  // per cell: the original index of each candidate (cand's layout)
    // rows whose appended part exceeds the fill's LDS column: their sorted
    // destination positions, and how many (sorted by k_varh_long_rows)
    int32_t *long_rows;
    int32_t *nlong;
    // per cell key: nonzero when a source cell lies within the stencil reach
    // of every destination (sph_cell_near), or NULL; the count pass then
    // skips a destination cell after two loads
    const uint8_t *s_near;
};

__device__ __forceinline__ int wave_max_i(int v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = max(v, __shfl_xor(v, d, 64));
    return v;
}

__device__ __forceinline__ int wave_incl_scan_i(int v)
{
    const int lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// broadcast lane `l` (wave-uniform) of a 64-bit value to the wavefront
__device__ __forceinline__ double wave_bcast(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ int64_t wave_bcast_i64(int64_t v, int l)
{
    const int lo = __builtin_amdgcn_readlane((int)(v & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane((int)(v >> 32), l);
    return ((int64_t)hi << 32) | (uint32_t)lo;
}

// lo[l] = xlo and hi[l] = xhi for the wave-uniform lane l and scalars xlo,
// xhi (v_writelane_b32 with the lane select in M0: one scalar operand per
// instruction besides M0 keeps within the constant bus; values and lane are
// written by scalar code).  M0 is saved and restored around the two writes.
__device__ __forceinline__ void wave_writelane2(uint32_t &lo, uint32_t &hi, uint32_t xlo,
                                                uint32_t xhi, int l)
{
    uint32_t m0_saved;
    asm volatile("s_mov_b32 %2, m0\n\ts_mov_b32 m0, %5\n\ts_nop 0\n\t"
                 "v_writelane_b32 %0, %3, m0\n\tv_writelane_b32 %1, %4, m0\n\t"
                 "s_mov_b32 m0, %2"
                 : "+v"(lo), "+v"(hi), "=&s"(m0_saved)
                 : "s"(xlo), "s"(xhi), "s"(l));
}

__device__ __forceinline__ uint64_t wave_bcast_u64(uint64_t v, int l)
{
    const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)v, l);
    const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}

// FILL = false: count pass.  Runs the distance tests, writes counts[t], the
//   64-bit hit mask of every (destination, candidate chunk) pair to a.masks,
//   the cell of every destination and the source of each candidate.  Cells
//   whose stencil needs more than one batch or more than NN_MAXCH chunks are
//   flagged in a.cell_ovf.
// FILL = true: writes the SELL-64 rows of the flagged cells only (the other
//   cells are written by k_nnps_fill_slices from the masks).
template <typename T> struct Quad;
template <> struct Quad<double> {
    using type = double4;
    static __device__ type make(double a, double b, double c, double d)
    {
        return make_double4(a, b, c, d);
    }
};
template <> struct Quad<float> {
    using type = float4;
    static __device__ type make(float a, float b, float c, float d)
    {
        return make_float4(a, b, c, d);
    }
};
template <typename T> __device__ __forceinline__ const T *drec_of(const CellArgs &a);
template <> __device__ __forceinline__ const double *drec_of<double>(const CellArgs &a)
{
    return a.drec;
}
template <> __device__ __forceinline__ const float *drec_of<float>(const CellArgs &a)
{
    return a.drec32;
}
template <typename T> __device__ __forceinline__ const T *srec_of(const CellArgs &a);
template <> __device__ __forceinline__ const double *srec_of<double>(const CellArgs &a)
{
    return a.srec;
}
template <> __device__ __forceinline__ const float *srec_of<float>(const CellArgs &a)
{
    return a.srec32;
}

// one destination cell w (one wavefront); seg_start, seg_end, dsts are the
// block's LDS arrays
template <bool FILL, typename T>
__device__ void nnps_cell(const CellArgs &a, int64_t w, int32_t *__restrict__ counts,
                          const int64_t *__restrict__ slice_off, int32_t *__restrict__ indices,
                          int32_t *seg_start, int32_t *seg_end, typename Quad<T>::type *dsts)
{
    const T *drec = drec_of<T>(a);
    const T *srec = srec_of<T>(a);
    const int lane = threadIdx.x;
    const int32_t key = a.dcells[w];
    if (key == a.trash) return;  // rows stay empty
    // no source cell within reach: every row of the cell stays empty (counts
    // start at 0) -- most fluid cells against a wall array
    if (!FILL && a.s_near && !a.s_near[key]) return;
    const int32_t ds = a.dstart[key], de = a.dstart[key + 1];
    const GridDev g =
        a.grids ? a.grids[a.key_sim ? a.key_sim[key] : a.sim_of[a.dperm[ds]]] : a.g;
    // the cell's coordinates, recomputed exactly as the binning did (from
    // the double coordinates it used)
    double c0x, c0y, c0z;
    if (a.ox) {
        const int32_t i0 = a.dperm[ds];
        c0x = a.ox[i0];
        c0y = a.oy[i0];
        c0z = a.oz[i0];
    } else {
        const T *r0 = drec + (int64_t)ds * SPH_REC;
        c0x = r0[SPH_R_X];
        c0y = r0[SPH_R_Y];
        c0z = r0[SPH_R_Z];
    }
    const int cx = min(max(cell_coord(c0x, g.cs) - g.cmin0, 0), g.d0 - 1);
    const int cy = min(max(cell_coord(c0y, g.cs) - g.cmin1, 0), g.d1 - 1);
    const int cz = min(max(cell_coord(c0z, g.cs) - g.cmin2, 0), g.d2 - 1);
    // stencil reach: the largest (int)ceil(h*rs/cell) over the cell; a
    // destination with a smaller reach only sees extra candidates that fail
    // its distance test, and its own stencil is a sub-box visited in order
    int dl = 0;
    for (int32_t t = ds + lane; t < de; t += CW) {
        double r = (double)drec[(int64_t)t * SPH_REC + SPH_R_H] * a.rs / g.cs;
        // float h can round up past the h the cell size came from: a reach
        // of 1 + 1e-7 cells would double the stencil for pairs that only
        // the float rounding could add
        if (sizeof(T) == sizeof(float)) r *= 1.0 - 0x1p-20;
        // bounded so that a non-finite h (rejected on the host, flagged in a
        // batch) cannot overflow the stencil arithmetic; a reach above 256
        // cells needs a cell size clamped below radius/256
        dl = max(dl, r > 0.0 ? (r < 256.0 ? (int)ceil(r) : 256) : 0);
    }
    const int dist = wave_max_i(dl);
    const int x0 = max(cx - dist, 0), x1 = min(cx + dist, g.d0 - 1);
    const int y0 = max(cy - dist, 0), y1 = min(cy + dist, g.d1 - 1);
    const int z0 = max(cz - dist, 0), z1 = min(cz + dist, g.d2 - 1);
    const int nx = x1 - x0 + 1, ny = y1 - y0 + 1, nz = z1 - z0 + 1;
    const int nst = nx * ny * nz;
    if (!FILL && !a.same && !a.s_near) {
        // no source in the stencil: every row of the cell stays empty (counts
        // start at 0) -- most fluid cells against a wall array.  (A cell's own
        // array always has its members in the stencil.)
        int32_t found = 0;
        for (int sIdx = lane; sIdx < nst; sIdx += CW) {
            const int iz = sIdx % nz, iy = (sIdx / nz) % ny, ix = sIdx / (nz * ny);
            const int32_t k = cell_key(g, x0 + ix, y0 + iy, z0 + iz);
            found |= a.sstart[k + 1] - a.sstart[k];
        }
        if (!__builtin_amdgcn_ballot_w64(found != 0)) return;
    }

    int nchunks = 0;  // chunks per destination over all stencil batches
    for (int32_t g0 = ds; g0 < de; g0 += CW) {  // destination groups of <= 64
        const int gn = min(CW, de - g0);
        const int32_t my_t = g0 + lane;  // destination owned by this lane
        bool mreal = false;
        if (lane < gn) {
            const T *dr = drec + (int64_t)my_t * SPH_REC;
            const T radius = dr[SPH_R_H] * T(a.rs);
            dsts[lane] = Quad<T>::make(dr[SPH_R_X], dr[SPH_R_Y], dr[SPH_R_Z], radius * radius);
            mreal = a.dperm[my_t] < a.nd_real;
        }
        const uint64_t realmask = __ballot(mreal);
        // SELL-64: entry k of row t lives at slice_off[t/64] + 64*k + t%64
        int64_t myaddr = 0;    // FILL: address of entry 0 of my row
        int32_t mycount = 0;   // entries found so far (COUNT and FILL)
        int32_t myself = -1;   // FILL: entry holding my own index
        int32_t mylast = -1;   // FILL: last neighbour written to my row
        if (FILL && lane < gn) myaddr = slice_off[my_t >> 6] + (my_t & 63);
        int cidx = 0;          // chunk index over the stencil batches
        for (int b0 = 0; b0 < nst; b0 += MAXSEG) {  // stencil batches
            const int sIdx = b0 + lane;
            int32_t st = 0, cnt = 0;
            if (lane < MAXSEG && sIdx < nst) {
                const int iz = sIdx % nz, iy = (sIdx / nz) % ny, ix = sIdx / (nz * ny);
                const int32_t k = cell_key(g, x0 + ix, y0 + iy, z0 + iz);
                st = a.sstart[k];
                cnt = a.sstart[k + 1] - st;
            }
            const int incl = wave_incl_scan_i(cnt);
            const int total = __shfl(incl, CW - 1, 64);
            __syncthreads();
            if (lane < MAXSEG) {
                seg_start[lane] = st;
                seg_end[lane] = incl;
            }
            __syncthreads();
            for (int c0 = 0; c0 < total; c0 += CW, ++cidx) {
                const int q = c0 + lane;
                const bool valid = q < total;
                int32_t p = 0;
                T sx = T(0.0), sy = T(0.0), sz = T(0.0);
                if (valid) {
                    int lo = 0, hi = MAXSEG - 1;  // first segment with end > q
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (seg_end[mid] > q) hi = mid; else lo = mid + 1;
                    }
                    const int prev = lo > 0 ? seg_end[lo - 1] : 0;
                    p = seg_start[lo] + (q - prev);
                    if (!FILL && g0 == ds && b0 == 0 && q < a.maxch * CW)
                        a.cand[w * (a.maxch * CW) + q] = p;
                    const T *sr = srec + (int64_t)p * SPH_REC;
                    sx = sr[SPH_R_X];
                    sy = sr[SPH_R_Y];
                    sz = sr[SPH_R_Z];
                }
                uint32_t mlo = 0, mhi = 0;  // count: my destination's hit mask of this chunk
                if (!FILL) {
                    // every lane runs this loop (EXEC is full): the hit mask
                    // of destination d is the ballot of the compare, kept in
                    // lane d with v_writelane (no per-lane select)
                    const uint64_t validm = __builtin_amdgcn_ballot_w64(valid);
                    for (int d = 0; d < gn; ++d) {
                        if (!((realmask >> d) & 1)) continue;  // non-LocalReal: empty row
                        const auto D = dsts[d];  // same address in every lane
                        // cPoint_distance2(src, dst), base/point.pxd:35-37
                        const T ex = sx - D.x, ey = sy - D.y, ez = sz - D.z;
                        const T d2 = ex * ex + ey * ey + ez * ez;
                        const uint64_t mask = __builtin_amdgcn_ballot_w64(d2 < D.w) & validm;
                        wave_writelane2(mlo, mhi, (uint32_t)mask, (uint32_t)(mask >> 32), d);
                    }
                }
                for (int d = 0; FILL && d < gn; ++d) {
                    if (!((realmask >> d) & 1)) continue;  // non-LocalReal: empty row
                    const int32_t t = g0 + d;
                    const auto D = dsts[d];  // same address in every lane
                    // cPoint_distance2(src, dst), base/point.pxd:35-37
                    const T ex = sx - D.x, ey = sy - D.y, ez = sz - D.z;
                    const T d2 = ex * ex + ey * ey + ez * ez;
                    const bool hit = valid && d2 < D.w;
                    const uint64_t mask = __ballot(hit);
                    if (mask == 0) continue;
                    {
                        const int nh = __popcll(mask);
                        const int64_t addr0 = wave_bcast_i64(myaddr, d);
                        const int32_t k0 = __builtin_amdgcn_readlane(mycount, d);
                        const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
                        if (hit)  // the slice's lines fill up in L2 from several cells
                            indices[addr0 + ((int64_t)(k0 + __popcll(mask & below)) << 6)] = p;
                        const int lastlane = 63 - __clzll(mask);
                        const int32_t lastp = __builtin_amdgcn_readlane(p, lastlane);
                        const uint64_t selfm = __ballot(a.same && hit && p == t);
                        if (lane == d) {
                            if (selfm) {
                                const int sl = __ffsll((long long)selfm) - 1;
                                const uint64_t bsl = sl == 0 ? 0ull : (~0ull >> (64 - sl));
                                myself = k0 + __popcll(mask & bsl);
                            }
                            mycount = k0 + nh;
                            mylast = lastp;
                        }
                    }
                }
                if (!FILL) {
                    mycount += __popc(mlo) + __popc(mhi);
                    if (cidx < a.maxch && lane < gn)
                        a.masks[(int64_t)cidx * a.mstride + my_t] = ((uint64_t)mhi << 32) | mlo;
                }
            }
        }
        nchunks = cidx;
        if (lane < gn) {
            if (!FILL) {
                counts[my_t] = mycount;
                a.dcell_of[my_t] = (int32_t)w;
            } else if (a.same && myself >= 0 && myself < mycount - 1) {
                // keep the particle itself at the last index (nnps.pyx:612-619)
                indices[myaddr + ((int64_t)myself << 6)] = mylast;
                indices[myaddr + ((int64_t)(mycount - 1) << 6)] = my_t;
            }
        }
    }
    if (!FILL && lane == 0) {
        const int ovf = nst > MAXSEG || nchunks > a.maxch;
        a.cell_ovf[w] = ovf ? -1 : nchunks;
        if (ovf) a.ovf_list[atomicAdd(a.novf, 1)] = (int32_t)w;
    }
}

// FILL = false: every non-empty destination cell (XCD x walks the x-th
// eighth of the cells with its blocks).  FILL = true: the cells the count
// pass flagged, listed in a.ovf_list.
// The count pass of one build over block bid of G: every non-empty destination
// cell (XCD x walks the x-th eighth of the cells with its blocks; bid is
// the block's XCD-consistent index, bid % 8 being its XCD).
template <typename T>
__device__ __forceinline__ void count_cells(const CellArgs &a, int32_t *__restrict__ counts,
                                            unsigned bid, unsigned G, int32_t *seg_start,
                                            int32_t *seg_end, typename Quad<T>::type *dsts)
{
    // cells read on the device are bounded by the launch size the scratch
    // was carved for (more cells than that is a status error of the caller,
    // whose lists are then incomplete, never written past)
    int64_t ncl = a.ncell_list;
    if (a.ncell_dev && (int64_t)*a.ncell_dev < ncl) ncl = *a.ncell_dev;
    // blocks go round-robin to the XCDs (block b to XCD b % 8): XCD x walks
    // cells [x*ncl/8, (x+1)*ncl/8) with its blocks, so every XCD has an
    // equal, contiguous share whatever the grid size (the grid is sized for a
    // bound when the cell count is read on the device)
    if (G < SPH_NXCD) {  // too few blocks to give every XCD one
        for (int64_t w = bid; w < ncl; w += G)
            nnps_cell<false, T>(a, w, counts, nullptr, nullptr, seg_start, seg_end, dsts);
        return;
    }
    const unsigned x = bid % SPH_NXCD, k = bid / SPH_NXCD;
    const unsigned per = (G - x + SPH_NXCD - 1) / SPH_NXCD;  // blocks on XCD x (>= 1)
    const int64_t lo = ncl * x / SPH_NXCD, hi = ncl * (x + 1) / SPH_NXCD;
    for (int64_t w = lo + k; w < hi; w += per)  // uniform: one wavefront per block
        nnps_cell<false, T>(a, w, counts, nullptr, nullptr, seg_start, seg_end, dsts);
}

// the rows of the cells the count pass flagged (listed in a.ovf_list), by
// blocks bid of G
template <typename T>
__device__ __forceinline__ void fill_flagged(const CellArgs &a, int32_t *__restrict__ counts,
                                             const int64_t *__restrict__ slice_off,
                                             int32_t *__restrict__ indices, unsigned bid,
                                             unsigned G, int32_t *seg_start, int32_t *seg_end,
                                             typename Quad<T>::type *dsts)
{
    if (a.status && slice_off[a.nslices] > a.cap) return;  // no room: nothing is written
    const int32_t nov = *a.novf;
    for (int32_t k = bid; k < nov; k += G)
        nnps_cell<true, T>(a, a.ovf_list[k], counts, slice_off, indices, seg_start, seg_end,
                           dsts);
}

// FILL = false: the count pass (count_cells).  FILL = true: the cells the
// count pass flagged (fill_flagged).
template <bool FILL, typename T>
__global__ void __launch_bounds__(CW)
    k_nnps_cells(CellArgs a, int32_t *__restrict__ counts, const int64_t *__restrict__ slice_off,
                 int32_t *__restrict__ indices)
{
    __shared__ int32_t seg_start[MAXSEG];
    __shared__ int32_t seg_end[MAXSEG];  // inclusive prefix of the segment sizes
    __shared__ typename Quad<T>::type dsts[CW];  // (x, y, z, radius^2) of the group's destinations
    if (FILL) {
        fill_flagged<T>(a, counts, slice_off, indices, blockIdx.x, gridDim.x, seg_start, seg_end,
                        dsts);
    } else {
        count_cells<T>(a, counts, blockIdx.x, gridDim.x, seg_start, seg_end, dsts);
    }
}

// The count passes of two builds in one launch: blocks [0, g0) the first
// (g0 a multiple of the XCD count, so block indices keep their XCD), the
// rest the second -- the smaller build fills the larger one's tail.
template <typename T>
__global__ void __launch_bounds__(CW)
    k_nnps_cells_pair(CellArgs a0, int32_t *__restrict__ counts0, CellArgs a1,
                      int32_t *__restrict__ counts1, unsigned g0)
{
    __shared__ int32_t seg_start[MAXSEG];
    __shared__ int32_t seg_end[MAXSEG];
    __shared__ typename Quad<T>::type dsts[CW];
    if (blockIdx.x < g0)
        count_cells<T>(a0, counts0, blockIdx.x, g0, seg_start, seg_end, dsts);
    else
        count_cells<T>(a1, counts1, blockIdx.x - g0, gridDim.x - g0, seg_start, seg_end, dsts);
}

// the flagged cells of two builds in one launch (blockIdx.y picks the build)
template <typename T>
__global__ void __launch_bounds__(CW)
    k_nnps_flagged_pair(CellArgs a0, int32_t *__restrict__ counts0,
                        const int64_t *__restrict__ slice_off0, int32_t *__restrict__ indices0,
                        CellArgs a1, int32_t *__restrict__ counts1,
                        const int64_t *__restrict__ slice_off1, int32_t *__restrict__ indices1)
{
    __shared__ int32_t seg_start[MAXSEG];
    __shared__ int32_t seg_end[MAXSEG];
    __shared__ typename Quad<T>::type dsts[CW];
    if (blockIdx.y)
        fill_flagged<T>(a1, counts1, slice_off1, indices1, blockIdx.x, gridDim.x, seg_start,
                        seg_end, dsts);
    else
        fill_flagged<T>(a0, counts0, slice_off0, indices0, blockIdx.x, gridDim.x, seg_start,
                        seg_end, dsts);
}

// SELL-64 rows from the count pass's hit masks: one lane per destination,
// one wavefront per 64-row slice, so that entry k of the 64 rows is one
// contiguous 256-byte store.  Candidate q = 64*chunk + bit of the cell's
// stencil is sorted source cand[cell][q].
// VARH: the rows of k_varh_cells -- the own part (fcount entries, from the
// own masks, own index not moved), then the appended part from a.amasks in
// ascending original index (the count pass's candj table), sorted in the
// lane's own column of the mask array once the own part no longer reads it
// (VARH_COL entries; a longer part is left to k_varh_long_rows) and written
// as sorted positions (srank).  Rows of cells k_varh_rows handles are
// skipped.
#define VARH_COL (2 * NN_MAXCH)
template <bool VARH>
__device__ __forceinline__ void fill_rows(const CellArgs &a, int64_t nd_total,
                                          int32_t *__restrict__ counts,
                                          const int64_t *__restrict__ slice_off,
                                          int32_t *__restrict__ indices,
                                          const int32_t *__restrict__ srank, unsigned bid,
                                          unsigned G, uint64_t (*s_mask)[NN_THREADS])
{
    if (a.totals && bid == 0 && threadIdx.x == 0) {
        // totals and capacity checks of a device-side build: status bit 0
        // when the lists need more than cap slots (nothing is written), bit
        // 4 when more cells are non-empty than the launch bound
        const int64_t sl = slice_off[a.nslices];
        int32_t st = *a.status;
        if (sl > a.cap) st |= 1;
        if (a.ncell_dev && (int64_t)*a.ncell_dev > a.cell_bound) st |= 16;
        *a.status = st;
        a.totals[0] = sl;
        a.totals[1] = *a.nnz_dev;
        a.totals[2] = st;
    }
    const int64_t t = (int64_t)sph_xcd_tile(bid, G) * NN_THREADS + threadIdx.x;
    if (t >= nd_total) return;
    if (a.status && slice_off[a.nslices] > a.cap) {
        counts[t] = 0;  // no room (status is set): empty rows, so no reader touches the slots
        return;
    }
    const int32_t nall = counts[t];
    if (nall == 0) return;
    const int32_t w = a.dcell_of[t];
    if (VARH && w < 0) return;
    const int32_t nch = a.cell_ovf[w];  // chunks of the cell's stencil
    if (nch < 0) return;                // written by the per-cell pass
    const int32_t n = VARH ? a.fcount[t] : nall;  // entries from the (own) masks
    int32_t *row = indices + slice_off[t >> 6] + (t & 63);
    const int32_t *cand = a.cand + (int64_t)w * (a.maxch * CW);
    const uint64_t *mk = a.masks + t;
    // the row's chunk masks, loaded together and kept in LDS (each lane reads
    // only its own column): walking them later waits on LDS, not on the
    // global stores of the row
    uint64_t v[NN_MAXCH];
#pragma unroll
    for (int c = 0; c < NN_MAXCH; ++c) v[c] = c < nch ? mk[(int64_t)c * a.mstride] : 0ull;
#pragma unroll
    for (int c = 0; c < NN_MAXCH; ++c) s_mask[c][threadIdx.x] = v[c];
    int c = 0;
    uint64_t m = v[0];
    int32_t self_k = -1, last_p = -1;
    // FILL_UNROLL candidates per round; the next round's table loads are
    // issued before this round's stores, so a round never waits for the
    // stores before it
    constexpr int FILL_UNROLL = 8;
    int32_t q[FILL_UNROLL], p[FILL_UNROLL];
    auto next_round = [&](int32_t k0, int32_t *qq, int32_t *pp) {
#pragma unroll
        for (int j = 0; j < FILL_UNROLL; ++j) {
            qq[j] = -1;
            if (k0 + j < n) {
                while (m == 0 && c + 1 < nch) m = s_mask[++c][threadIdx.x];
                if (m != 0) {  // always: the masks hold exactly n bits
                    qq[j] = (c << 6) + __ffsll((unsigned long long)m) - 1;
                    m &= m - 1;
                }
            }
        }
        // unconditional loads (entry 0 of the cell's table stands in for an
        // empty slot): straight-line code, so each store waits only for its
        // own round's loads
#pragma unroll
        for (int j = 0; j < FILL_UNROLL; ++j) pp[j] = cand[qq[j] >= 0 ? qq[j] : 0];
    };
    auto store_round = [&](int32_t k0, const int32_t *qq, const int32_t *pp) {
#pragma unroll
        for (int j = 0; j < FILL_UNROLL; ++j) {
            if (qq[j] < 0) continue;
            row[(int64_t)(k0 + j) << 6] = pp[j];
            if (pp[j] == t) self_k = k0 + j;
            last_p = pp[j];
        }
    };
    // two rounds per iteration, alternating buffers (no register copies
    // between rounds)
    int32_t qn[FILL_UNROLL], pn[FILL_UNROLL];
    next_round(0, q, p);
    for (int32_t k0 = 0; k0 < n; k0 += 2 * FILL_UNROLL) {
        next_round(k0 + FILL_UNROLL, qn, pn);
        store_round(k0, q, p);
        if (k0 + FILL_UNROLL >= n) break;
        next_round(k0 + 2 * FILL_UNROLL, q, p);
        store_round(k0 + FILL_UNROLL, qn, pn);
    }
    if (!VARH) {
        if (a.same && self_k >= 0 && self_k < n - 1) {
            // keep the particle itself at the last index (nnps.pyx:612-619)
            row[(int64_t)self_k << 6] = last_p;
            row[(int64_t)(n - 1) << 6] = (int32_t)t;
        }
        return;
    }
    // the appended part: real sources outside the own radius that reach t,
    // ascending original index (base/nnps.pyx:742-786)
    const int32_t ne = nall - n;
    if (ne <= 0) return;
    // the lane's column of s_mask as 32-bit slots (only this lane uses them)
    uint32_t *col = reinterpret_cast<uint32_t *>(&s_mask[0][0]) + 2 * threadIdx.x;
    auto slot = [&](int32_t k) -> uint32_t & {
        return col[(int64_t)(k >> 1) * (2 * NN_THREADS) + (k & 1)];
    };
    const int32_t *candj = a.candj + (int64_t)w * (a.maxch * CW);
    if (ne <= VARH_COL) {
        // the chunk masks loaded together; the candidate numbers into LDS
        // (no loads), then their original indices by loads four at a time,
        // sorted, turned back into sorted positions the same way and stored:
        // no load waits behind a store or a single other load of this part
        uint64_t am[NN_MAXCH];
#pragma unroll
        for (int c2 = 0; c2 < NN_MAXCH; ++c2)
            am[c2] = c2 < nch ? a.amasks[(int64_t)c2 * a.mstride + t] : 0ull;
        int32_t e = 0;
#pragma unroll
        for (int c2 = 0; c2 < NN_MAXCH; ++c2) {
            uint64_t m2 = am[c2];
            while (m2 != 0) {
                const int b = __ffsll((unsigned long long)m2) - 1;
                m2 &= m2 - 1;
                if (e < ne) slot(e) = (uint32_t)((c2 << 6) + b);
                ++e;
            }
        }
        const int32_t na = min(e, ne);
        {
            int32_t i = 0;
            for (; i + 4 <= na; i += 4) {
                const int32_t j0 = candj[slot(i)], j1 = candj[slot(i + 1)];
                const int32_t j2 = candj[slot(i + 2)], j3 = candj[slot(i + 3)];
                slot(i) = (uint32_t)j0;
                slot(i + 1) = (uint32_t)j1;
                slot(i + 2) = (uint32_t)j2;
                slot(i + 3) = (uint32_t)j3;
            }
            for (; i < na; ++i) slot(i) = (uint32_t)candj[slot(i)];
        }
        for (int32_t i = 1; i < na; ++i) {
            const uint32_t v = slot(i);
            int32_t q2 = i - 1;
            while (q2 >= 0 && slot(q2) > v) {
                slot(q2 + 1) = slot(q2);
                --q2;
            }
            slot(q2 + 1) = v;
        }
        int32_t i = 0;
        for (; i + 4 <= na; i += 4) {
            const int32_t r0 = srank[slot(i)], r1 = srank[slot(i + 1)];
            const int32_t r2 = srank[slot(i + 2)], r3 = srank[slot(i + 3)];
            slot(i) = (uint32_t)r0;
            slot(i + 1) = (uint32_t)r1;
            slot(i + 2) = (uint32_t)r2;
            slot(i + 3) = (uint32_t)r3;
        }
        for (; i < na; ++i) slot(i) = (uint32_t)srank[slot(i)];
        for (i = 0; i < na; ++i) row[(int64_t)(n + i) << 6] = (int32_t)slot(i);
        return;
    }
    // a long appended part: its original indices written unsorted, the row
    // listed for k_varh_long_rows (one wavefront sorts it in LDS)
    int32_t e = 0;
    for (int c2 = 0; c2 < nch; ++c2) {
        uint64_t am = a.amasks[(int64_t)c2 * a.mstride + t];
        while (am != 0) {
            const int b = __ffsll((unsigned long long)am) - 1;
            am &= am - 1;
            if (e < ne) row[(int64_t)(n + e) << 6] = candj[(c2 << 6) + b];
            ++e;
        }
    }
    a.long_rows[atomicAdd(a.nlong, 1)] = (int32_t)t;
}

template <bool VARH>
__global__ void __launch_bounds__(NN_THREADS)
    k_nnps_fill_slices(CellArgs a, int64_t nd_total, int32_t *__restrict__ counts,
                       const int64_t *__restrict__ slice_off, int32_t *__restrict__ indices,
                       const int32_t *__restrict__ srank)
{
    __shared__ uint64_t s_mask[NN_MAXCH][NN_THREADS];
    fill_rows<VARH>(a, nd_total, counts, slice_off, indices, srank, blockIdx.x, gridDim.x,
                    s_mask);
}

// The fills of two fixed-h builds in one launch (blocks [0, g0) the first, g0
// a multiple of the XCD count)
__global__ void __launch_bounds__(NN_THREADS)
    k_nnps_fill_pair(CellArgs a0, int64_t nd0, int32_t *__restrict__ counts0,
                     const int64_t *__restrict__ slice_off0, int32_t *__restrict__ indices0,
                     CellArgs a1, int64_t nd1, int32_t *__restrict__ counts1,
                     const int64_t *__restrict__ slice_off1, int32_t *__restrict__ indices1,
                     unsigned g0)
{
    __shared__ uint64_t s_mask[NN_MAXCH][NN_THREADS];
    if (blockIdx.x < g0)
        fill_rows<false>(a0, nd0, counts0, slice_off0, indices0, nullptr, blockIdx.x, g0, s_mask);
    else
        fill_rows<false>(a1, nd1, counts1, slice_off1, indices1, nullptr, blockIdx.x - g0,
                         gridDim.x - g0, s_mask);
}

// The appended parts the fill left unsorted (longer than its LDS column): one
// wavefront per listed row sorts the original indices ascending (bitonic, in
// LDS, up to VARH_LONG entries; a longer part by one lane in place) and
// writes them back as sorted positions (srank).
#define VARH_LONG 1024
__global__ void __launch_bounds__(CW)
    k_varh_long_rows(CellArgs a, const int32_t *__restrict__ counts,
                     const int64_t *__restrict__ slice_off, int32_t *__restrict__ indices,
                     const int32_t *__restrict__ srank)
{
    __shared__ int32_t s_v[VARH_LONG];
    const int lane = threadIdx.x;
    const int32_t nl = *a.nlong;
    for (int32_t r = blockIdx.x; r < nl; r += gridDim.x) {
        const int32_t t = a.long_rows[r];
        const int32_t n = a.fcount[t];
        const int32_t ne = counts[t] - n;
        int32_t *app = indices + slice_off[t >> 6] + (t & 63) + ((int64_t)n << 6);
        if (ne > VARH_LONG) {  // one lane, in place
            if (lane == 0) {
                for (int32_t i = 1; i < ne; ++i) {
                    const int32_t v = app[(int64_t)i << 6];
                    int32_t q2 = i - 1;
                    while (q2 >= 0 && app[(int64_t)q2 << 6] > v) {
                        app[(int64_t)(q2 + 1) << 6] = app[(int64_t)q2 << 6];
                        --q2;
                    }
                    app[(int64_t)(q2 + 1) << 6] = v;
                }
                for (int32_t i = 0; i < ne; ++i) app[(int64_t)i << 6] = srank[app[(int64_t)i << 6]];
            }
            continue;
        }
        int32_t P = CW;
        while (P < ne) P <<= 1;
        for (int32_t i = lane; i < P; i += CW) s_v[i] = i < ne ? app[(int64_t)i << 6] : INT32_MAX;
        __syncthreads();
        for (int32_t k = 2; k <= P; k <<= 1)
            for (int32_t j = k >> 1; j > 0; j >>= 1) {
                for (int32_t i = lane; i < P; i += CW) {
                    const int32_t ixj = i ^ j;
                    if (ixj > i) {
                        const int32_t x = s_v[i], y = s_v[ixj];
                        if ((x > y) == ((i & k) == 0)) {
                            s_v[i] = y;
                            s_v[ixj] = x;
                        }
                    }
                }
                __syncthreads();
            }
        for (int32_t i = lane; i < ne; i += CW) app[(int64_t)i << 6] = srank[s_v[i]];
        __syncthreads();  // s_v is reused by the next row
    }
}

// slice s of 64 destinations: storage 64 * max(count) and its entry total;
// the storage also in slice_w2[s] (read by k_slice_offsets while slice_w,
// which becomes the offsets, is rewritten)
struct SliceArgs {
    const int32_t *counts;
    int64_t nd_total, nslices;
    int64_t *slice_w, *slice_sum, *slice_w2;
};

__device__ __forceinline__ void slices_block(const SliceArgs &a, int64_t bid)
{
    const int64_t t = bid * NN_THREADS + threadIdx.x;
    const int c = t < a.nd_total ? a.counts[t] : 0;
    int m = c, sum = c;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        m = max(m, __shfl_xor(m, d, 64));
        sum += __shfl_xor(sum, d, 64);
    }
    const int64_t s = t >> 6;
    if ((threadIdx.x & 63) == 0 && s < a.nslices) {
        a.slice_w[s] = 64 * (int64_t)m;
        a.slice_w2[s] = 64 * (int64_t)m;
        a.slice_sum[s] = sum;
    }
}

__global__ void __launch_bounds__(NN_THREADS)
    k_slices(const int32_t *__restrict__ counts, int64_t nd_total, int64_t nslices,
             int64_t *__restrict__ slice_w, int64_t *__restrict__ slice_sum,
             int64_t *__restrict__ slice_w2)
{
    const SliceArgs a = {counts, nd_total, nslices, slice_w, slice_sum, slice_w2};
    slices_block(a, blockIdx.x);
}

// two builds' slices in one launch: blockIdx.y picks the build, blocks past
// its own destinations have nothing to do
__global__ void __launch_bounds__(NN_THREADS) k_slices_pair(SliceArgs a0, SliceArgs a1)
{
    const SliceArgs &a = blockIdx.y ? a1 : a0;
    if ((int64_t)blockIdx.x * NN_THREADS >= a.nslices * 64) return;
    slices_block(a, blockIdx.x);
}

// counts of every row and the flagged-cell counter start at 0
__global__ void __launch_bounds__(NN_THREADS)
    k_nnps_reset(int32_t *__restrict__ counts, int64_t nd_total, int32_t *__restrict__ novf)
{
    const int64_t t = (int64_t)blockIdx.x * NN_THREADS + threadIdx.x;
    if (t < nd_total) counts[t] = 0;
    if (t == 0) *novf = 0;
}

// the same for two builds (blockIdx.y picks the build)
__global__ void __launch_bounds__(NN_THREADS)
    k_nnps_reset_pair(int32_t *__restrict__ counts0, int64_t nd0, int32_t *__restrict__ novf0,
                      int32_t *__restrict__ counts1, int64_t nd1, int32_t *__restrict__ novf1)
{
    int32_t *counts = blockIdx.y ? counts1 : counts0;
    int32_t *novf = blockIdx.y ? novf1 : novf0;
    const int64_t nd = blockIdx.y ? nd1 : nd0;
    const int64_t t = (int64_t)blockIdx.x * NN_THREADS + threadIdx.x;
    if (t < nd) counts[t] = 0;
    if (t == 0) *novf = 0;
}

// total of n int64 values by one block (deterministic tree order)
__global__ void __launch_bounds__(1024)
    k_sum_i64(const int64_t *__restrict__ v, int64_t n, int64_t *__restrict__ out)
{
    __shared__ int64_t part[1024 / 64];
    int64_t s = 0;
    for (int64_t i = threadIdx.x; i < n; i += 1024) s += v[i];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t tot = 0;
        for (int w = 0; w < 1024 / 64; ++w) tot += part[w];
        *out = tot;
    }
}

// The slice widths (in slice_off[0..nslices), a copy in width) turned into
// offsets and the entry total in two launches when there are few tiles of
// slices: per-tile sums of the widths and entry counts, then every tile adds
// the sums of the tiles before it to its own scan and writes slice_off (the
// last tile also slice_off[nslices] and *nnz).  Every load of a thread is
// issued before the first is used.  (The general path, k_sum_i64 and a
// multi-level scan, takes four launches.)
#define SLICE_THREADS 256
#define SLICE_ITEMS 16
#define SLICE_TILE (SLICE_THREADS * SLICE_ITEMS)
#define SLICE_MAX_TILES 256

__device__ __forceinline__ void slice_block_sum2(int64_t &a, int64_t &b, longlong2 *s_part)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        a += __shfl_xor(a, d, 64);
        b += __shfl_xor(b, d, 64);
    }
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = make_longlong2(a, b);
    __syncthreads();
    a = b = 0;
    for (int q = 0; q < SLICE_THREADS / 64; ++q) {
        a += s_part[q].x;
        b += s_part[q].y;
    }
    __syncthreads();
}

// item k of thread t of tile b: slice b*TILE + t + k*THREADS (coalesced)
__device__ __forceinline__ void slice_tile_sums(const int64_t *__restrict__ width,
                                                const int64_t *__restrict__ slice_sum,
                                                int64_t nslices,
                                                longlong2 *__restrict__ tile_sum, int64_t bid,
                                                longlong2 *s_part)
{
    const int64_t base = bid * SLICE_TILE + threadIdx.x;
    const int64_t last = nslices - 1;
    int64_t w[SLICE_ITEMS], z[SLICE_ITEMS];
#pragma unroll
    for (int k = 0; k < SLICE_ITEMS; ++k) {
        const int64_t q = base + (int64_t)k * SLICE_THREADS;
        w[k] = width[q < last ? q : last];
        z[k] = slice_sum[q < last ? q : last];
    }
    int64_t a = 0, b = 0;
#pragma unroll
    for (int k = 0; k < SLICE_ITEMS; ++k) {
        const bool in = base + (int64_t)k * SLICE_THREADS < nslices;
        a += in ? w[k] : 0;
        b += in ? z[k] : 0;
    }
    slice_block_sum2(a, b, s_part);
    if (threadIdx.x == 0) tile_sum[bid] = make_longlong2(a, b);
}

__global__ void __launch_bounds__(SLICE_THREADS)
    k_slice_tile_sums(const int64_t *__restrict__ width, const int64_t *__restrict__ slice_sum,
                      int64_t nslices, longlong2 *__restrict__ tile_sum)
{
    __shared__ longlong2 s_part[SLICE_THREADS / 64];
    slice_tile_sums(width, slice_sum, nslices, tile_sum, blockIdx.x, s_part);
}

// tile bid of ntiles: its slice offsets, and (last tile) the totals
__device__ __forceinline__ void slice_tile_scan(const int64_t *__restrict__ width,
                                                int64_t nslices,
                                                const longlong2 *__restrict__ tile_sum,
                                                int64_t *__restrict__ slice_off,
                                                int64_t *__restrict__ nnz, int64_t bid,
                                                int64_t ntiles, longlong2 *s_part,
                                                int64_t (*s_round)[SLICE_THREADS / 64])
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t base = bid * SLICE_TILE + tid;
    const int64_t last = nslices - 1;
    int64_t w[SLICE_ITEMS];
#pragma unroll
    for (int k = 0; k < SLICE_ITEMS; ++k) {
        const int64_t q = base + (int64_t)k * SLICE_THREADS;
        w[k] = q < nslices ? width[q < last ? q : last] : 0;
    }
    // sums of the tiles before this one (and of all tiles, for the totals)
    int64_t pa = 0, pz = 0, ta = 0, tz = 0;
    for (int64_t q = tid; q < ntiles; q += SLICE_THREADS) {
        const longlong2 v = tile_sum[q];
        if (q < bid) {
            pa += v.x;
        }
        ta += v.x;
        tz += v.y;
    }
    slice_block_sum2(pa, pz, s_part);
    slice_block_sum2(ta, tz, s_part);
    // round k: a scan over the block of the items k (THREADS consecutive slices)
    int64_t inc[SLICE_ITEMS];
#pragma unroll
    for (int k = 0; k < SLICE_ITEMS; ++k) {
        int64_t x = w[k];
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int64_t up = __shfl_up(x, d, 64);
            if (lane >= d) x += up;
        }
        inc[k] = x;
        if (lane == 63) s_round[k][wave] = x;
    }
    __syncthreads();
    int64_t carry = pa;
#pragma unroll
    for (int k = 0; k < SLICE_ITEMS; ++k) {
        int64_t before = 0, total = 0;
#pragma unroll
        for (int q = 0; q < SLICE_THREADS / 64; ++q) {
            const int64_t v = s_round[k][q];
            before += q < wave ? v : 0;
            total += v;
        }
        const int64_t idx = base + (int64_t)k * SLICE_THREADS;
        if (idx < nslices) slice_off[idx] = carry + before + inc[k] - w[k];
        carry += total;
    }
    if (bid == ntiles - 1 && tid == 0) {
        slice_off[nslices] = ta;
        *nnz = tz;
    }
}

__global__ void __launch_bounds__(SLICE_THREADS)
    k_slice_tile_scan(const int64_t *__restrict__ width, int64_t nslices,
                      const longlong2 *__restrict__ tile_sum, int64_t *__restrict__ slice_off,
                      int64_t *__restrict__ nnz)
{
    __shared__ longlong2 s_part[SLICE_THREADS / 64];
    __shared__ int64_t s_round[SLICE_ITEMS][SLICE_THREADS / 64];
    slice_tile_scan(width, nslices, tile_sum, slice_off, nnz, blockIdx.x, gridDim.x, s_part,
                    s_round);
}

// the two launches for two builds (blockIdx.y picks the build; a build's
// tiles are its first blocks, the rest have nothing to do)
struct SliceScan {
    const int64_t *width, *slice_sum;
    int64_t nslices, ntiles;
    longlong2 *tile_sum;
    int64_t *slice_off, *nnz;
};

__global__ void __launch_bounds__(SLICE_THREADS) k_slice_tile_sums_pair(SliceScan a0, SliceScan a1)
{
    __shared__ longlong2 s_part[SLICE_THREADS / 64];
    const SliceScan &a = blockIdx.y ? a1 : a0;
    if ((int64_t)blockIdx.x >= a.ntiles) return;
    slice_tile_sums(a.width, a.slice_sum, a.nslices, a.tile_sum, blockIdx.x, s_part);
}

__global__ void __launch_bounds__(SLICE_THREADS) k_slice_tile_scan_pair(SliceScan a0, SliceScan a1)
{
    __shared__ longlong2 s_part[SLICE_THREADS / 64];
    __shared__ int64_t s_round[SLICE_ITEMS][SLICE_THREADS / 64];
    const SliceScan &a = blockIdx.y ? a1 : a0;
    if ((int64_t)blockIdx.x >= a.ntiles) return;
    slice_tile_scan(a.width, a.nslices, a.tile_sum, a.slice_off, a.nnz, blockIdx.x, a.ntiles,
                    s_part, s_round);
}

// slice offsets and entry total from the widths k_slices wrote into
// slice_off (and slice_w2) and its per-slice sums
// The sums region of a build: per-slice entry totals (nslices), the entry
// total, a copy of the widths (nslices), then the tile sums (16 B each)
static size_t slice_sums_bytes(int64_t nslices)
{
    const int64_t ntiles = (nslices + SLICE_TILE - 1) / SLICE_TILE;
    return (((size_t)(2 * nslices + 2 + 2 * ntiles) * sizeof(int64_t) + 256) + 255) &
           ~(size_t)255;
}

static int slices_to_offsets(sph_ctx *ctx, int64_t *slice_off, const int64_t *slice_w2,
                             const int64_t *slice_sum, int64_t nslices, int64_t *nnz_dev,
                             size_t scan_scratch_off)
{
    const int64_t ntiles = (nslices + SLICE_TILE - 1) / SLICE_TILE;
    if (nslices > 0 && ntiles <= SLICE_MAX_TILES) {
        // after the copy of the widths (16-byte aligned: slice_sum is)
        longlong2 *tile_sum = (longlong2 *)(const_cast<int64_t *>(slice_sum) + 2 * nslices + 2);
        hipLaunchKernelGGL(k_slice_tile_sums, dim3((unsigned)ntiles), dim3(SLICE_THREADS), 0,
                           ctx->stream, slice_w2, slice_sum, nslices, tile_sum);
        SPH_LAUNCH_CHECK(ctx, "k_slice_tile_sums");
        hipLaunchKernelGGL(k_slice_tile_scan, dim3((unsigned)ntiles), dim3(SLICE_THREADS), 0,
                           ctx->stream, slice_w2, nslices, (const longlong2 *)tile_sum,
                           slice_off, nnz_dev);
        SPH_LAUNCH_CHECK(ctx, "k_slice_tile_scan");
        return SPH_OK;
    }
    // deterministic one-block total (atomics on one word from every slice
    // serialise: 190 us for 15,625 slices)
    hipLaunchKernelGGL(k_sum_i64, dim3(1), dim3(1024), 0, ctx->stream, slice_sum, nslices,
                       nnz_dev);
    SPH_LAUNCH_CHECK(ctx, "k_sum_i64");
    return sph_scan_exclusive_i64_at(ctx, slice_off, nslices, scan_scratch_off);
}

// status / totals of a device-side build: bit 0 of *status when the lists
// need more than cap slots; *nnz_acc += nnz (pair count kept on the device)
__global__ void k_nnps_dev_status(const int64_t *__restrict__ slots, const int64_t *__restrict__ nnz,
                                  int64_t cap, int32_t *__restrict__ status,
                                  int64_t *__restrict__ nnz_acc)
{
    if (threadIdx.x != 0) return;
    if (*slots > cap) *status |= 1;
    if (nnz_acc) *nnz_acc += *nnz;
}

// Count pass, slice offsets and totals, then (after the caller's check of
// the capacity) the fill.  Scratch: [hit masks | cell flags | candidate
// sources | cell of destination | slice sums | scan partials].
// Scratch of one build's count pass, from base (bytes, 256-aligned
// regions): hit masks | flagged cells | candidate tables | cell of each
// destination | flagged-cell list | slice sums | scan partials
struct CountScratch {
    size_t masks, ovf, cand, dcell, list, sums, scan, total;
};

static CountScratch count_scratch(const CellArgs &a, int64_t ncell_launch, int64_t nd_total)
{
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const int64_t nslices = (nd_total + 63) / 64;
    CountScratch c;
    c.masks = al((size_t)a.maxch * nd_total * sizeof(uint64_t));
    c.ovf = al((size_t)ncell_launch * sizeof(int32_t));
    c.cand = al((size_t)ncell_launch * a.maxch * CW * sizeof(int32_t));
    c.dcell = al((size_t)nd_total * sizeof(int32_t));
    c.list = al((size_t)(ncell_launch + 1) * sizeof(int32_t));
    c.sums = slice_sums_bytes(nslices);
    c.scan = al(sph_scan_scratch_bytes(nslices, sizeof(int64_t)));
    c.total = c.masks + c.ovf + c.cand + c.dcell + c.list + c.sums + c.scan;
    return c;
}

// the count pass's pointers in the scratch at offset base; the slice sums
// and the scan partials' offset for the slice offsets
static void count_carve(sph_ctx *ctx, CellArgs &a, const CountScratch &c, size_t base,
                        int64_t nd_total, int64_t **slice_sum, size_t *scan_off)
{
    char *sc = (char *)ctx->scratch + base;
    a.masks = (uint64_t *)sc;
    a.mstride = nd_total;
    a.cell_ovf = (int32_t *)(sc + c.masks);
    a.cand = (int32_t *)(sc + c.masks + c.ovf);
    a.dcell_of = (int32_t *)(sc + c.masks + c.ovf + c.cand);
    a.novf = (int32_t *)(sc + c.masks + c.ovf + c.cand + c.dcell);
    a.ovf_list = a.novf + 1;
    a.nslices = (nd_total + 63) / 64;
    const size_t sums_off = c.masks + c.ovf + c.cand + c.dcell + c.list;
    *slice_sum = (int64_t *)(sc + sums_off);
    *scan_off = base + sums_off + c.sums;
}

static int count_reset(sph_ctx *ctx, const CellArgs &a, int32_t *counts, int64_t nd_total)
{
    hipLaunchKernelGGL(k_nnps_reset, dim3(sph_blocks(nd_total, NN_THREADS)), dim3(NN_THREADS), 0,
                       ctx->stream, counts, nd_total, a.novf);
    SPH_LAUNCH_CHECK(ctx, "k_nnps_reset");
    return SPH_OK;
}

// slice widths -> offsets (scan); slice totals -> entry total
static int count_slices(sph_ctx *ctx, const CellArgs &a, const int32_t *counts,
                        int64_t *slice_off, int64_t *slice_sum, size_t scan_off,
                        int64_t nd_total)
{
    const int64_t nslices = a.nslices;
    hipLaunchKernelGGL(k_slices, dim3(sph_blocks(nslices * 64, NN_THREADS)), dim3(NN_THREADS), 0,
                       ctx->stream, counts, nd_total, nslices, slice_off, slice_sum,
                       slice_sum + nslices + 1);
    SPH_LAUNCH_CHECK(ctx, "k_slices");
    return slices_to_offsets(ctx, slice_off, slice_sum + nslices + 1, slice_sum, nslices,
                             slice_sum + nslices, scan_off);
}

static int64_t count_grid(const CellArgs &a, int64_t ncell_launch)
{
    const int cpb = a.cells_per_block > 0 ? a.cells_per_block : 1;
    const int64_t cgrid = (ncell_launch + cpb - 1) / cpb;
    return cgrid < NN_CELL_GRID ? cgrid : NN_CELL_GRID;
}

static int nnps_count(sph_ctx *ctx, CellArgs &a, int64_t ncell_launch, int64_t nd_total,
                      int32_t *counts, int64_t *slice_off, int64_t **nnz_dev)
{
    const CountScratch c = count_scratch(a, ncell_launch, nd_total);
    int rc = sph_ensure_scratch(ctx, c.total);
    if (rc) return rc;
    int64_t *slice_sum;
    size_t scan_off;
    count_carve(ctx, a, c, 0, nd_total, &slice_sum, &scan_off);
    *nnz_dev = slice_sum + a.nslices;
    rc = count_reset(ctx, a, counts, nd_total);
    if (rc) return rc;
    auto count_kern = a.f32 ? k_nnps_cells<false, float> : k_nnps_cells<false, double>;
    hipLaunchKernelGGL(count_kern, dim3((unsigned)count_grid(a, ncell_launch)), dim3(CW), 0,
                       ctx->stream, a, counts, (const int64_t *)nullptr, (int32_t *)nullptr);
    SPH_LAUNCH_CHECK(ctx, "k_nnps_cells<count>");
    return count_slices(ctx, a, counts, slice_off, slice_sum, scan_off, nd_total);
}

static int nnps_fill(sph_ctx *ctx, const CellArgs &a, int64_t ncell_launch, int64_t nd_total,
                     int32_t *counts, const int64_t *slice_off, int32_t *indices)
{
    hipLaunchKernelGGL(k_nnps_fill_slices<false>, dim3(sph_blocks(nd_total, NN_THREADS)),
                       dim3(NN_THREADS), 0, ctx->stream, a, nd_total, counts, slice_off, indices,
                       (const int32_t *)nullptr);
    SPH_LAUNCH_CHECK(ctx, "k_nnps_fill_slices");
    auto fill_kern = a.f32 ? k_nnps_cells<true, float> : k_nnps_cells<true, double>;
    const int64_t fill_grid = ncell_launch < NN_OVF_GRID ? ncell_launch : NN_OVF_GRID;
    hipLaunchKernelGGL(fill_kern, dim3((unsigned)fill_grid), dim3(CW), 0, ctx->stream, a,
                       (int32_t *)counts, slice_off, indices);
    SPH_LAUNCH_CHECK(ctx, "k_nnps_cells<fill>");
    return SPH_OK;
}

static int nnps_clear(sph_ctx *ctx, int64_t nd_total, int32_t *counts, int64_t *slice_off)
{
    const int64_t nslices = (nd_total + 63) / 64;
    // every row starts empty: non-real rows and cells without candidates stay 0
    hipLaunchKernelGGL(k_fill_i32, dim3(sph_blocks(nd_total + 1, NN_THREADS)), dim3(NN_THREADS),
                       0, ctx->stream, counts, nd_total, 0);
    SPH_LAUNCH_CHECK(ctx, "k_fill_i32");
    hipLaunchKernelGGL(k_fill_i32, dim3(sph_blocks(2 * (nslices + 1), NN_THREADS)),
                       dim3(NN_THREADS), 0, ctx->stream, (int32_t *)slice_off,
                       2 * (nslices + 1), 0);
    SPH_LAUNCH_CHECK(ctx, "k_fill_i32");
    return SPH_OK;
}

static CellArgs cell_args_base()
{
    CellArgs a;
    memset(&a, 0, sizeof(a));
    a.maxch = NN_MAXCH;
    a.trash = -1;
    return a;
}

// totals of an empty device-side plain build: totals = {slots, nnz, status}; status bit 0
// when the lists need more than cap slots (nothing written, rows empty), bit
// 4 when there are more non-empty destination cells than max_cells (the
// cells past that bound get no rows)
__global__ void k_nnps_dev_totals(const int64_t *__restrict__ slots,
                                  const int64_t *__restrict__ nnz,
                                  const int32_t *__restrict__ ncells_dev, int64_t cap,
                                  int64_t max_cells, int64_t *__restrict__ totals,
                                  int32_t *__restrict__ status)
{
    if (threadIdx.x != 0) return;
    const int64_t sl = slots ? *slots : 0, nz = nnz ? *nnz : 0;
    int32_t st = *status;
    if (sl > cap) st |= 1;
    if (ncells_dev && *ncells_dev > max_cells) st |= 16;
    *status = st;
    totals[0] = sl;
    totals[1] = nz;
    totals[2] = st;
}

// Host mode (dev == NULL): nd_cells is the host count of non-empty cells;
// the call reads the totals back (synchronises) and returns SPH_ENEEDCAP
// without writing lists when cap is too small.  Device mode: the cell count
// is read on the device (nd_cells bounds it), the totals go to dev->totals,
// the capacity checks to dev->status, and nothing synchronises.
struct NnpsDevOut {
    const int32_t *ncells_dev;
    int64_t *totals;
    int32_t *status;
};

// One build's validated arguments (nnps_prepare).  empty: every row was
// cleared (and a device-side build's totals written), nothing more to launch.
struct NnpsJob {
    CellArgs a;
    int64_t nd_cells, nd_total;
    int32_t *counts;
    int64_t *slice_off;
    int32_t *indices;
    bool empty;
};

static int nnps_prepare(sph_ctx *ctx, const sph_grid *grid, int f32, const double *d_rec,
                        const float *d_rec32, const double *ox, const double *oy,
                        const double *oz, const int32_t *d_perm, const int32_t *d_cell_start,
                        const int32_t *d_cells, int64_t nd_cells, int64_t nd_total,
                        int64_t nd_real, const int32_t *s_cell_start, const double *s_rec,
                        const float *s_rec32, int64_t ns, double radius_scale, int same_array,
                        int32_t *counts, int64_t *slice_off, int32_t *indices, int64_t cap,
                        int64_t *slots_out, int64_t *nnz_out, const NnpsDevOut *dev,
                        const uint8_t *s_near, int near_reach, NnpsJob *job)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, !s_near || (near_reach >= 1 && near_reach <= 8),
                "sph_nnps_build: near_reach must be in [1, 8] with a near-cells table");
    SPH_REQUIRE(ctx, grid && counts && slice_off &&
                         (dev ? (dev->ncells_dev && dev->totals && dev->status && indices)
                              : (slots_out && nnz_out)),
                "sph_nnps_build: null argument");
    SPH_REQUIRE(ctx, nd_total >= 0 && nd_total < ((int64_t)1 << 31) && nd_real >= 0 &&
                         nd_real <= nd_total && nd_cells >= 0 && nd_cells <= nd_total,
                "sph_nnps_build: bad destination counts");
    SPH_REQUIRE(ctx, ns == 0 || (s_cell_start && (f32 ? (const void *)s_rec32 : s_rec)),
                "sph_nnps_build: null source arrays");
    SPH_REQUIRE(ctx, nd_total == 0 || ((f32 ? (const void *)d_rec32 : d_rec) && d_perm &&
                                       d_cell_start && d_cells && (!f32 || (ox && oy && oz))),
                "sph_nnps_build: null destination arrays");
    int rc = sph_upload_key_tables(ctx, grid);
    if (rc) return rc;
    if (!dev) {
        *slots_out = 0;
        *nnz_out = 0;
    }
    job->nd_cells = nd_cells;
    job->nd_total = nd_total;
    job->counts = counts;
    job->slice_off = slice_off;
    job->indices = indices;
    job->empty = nd_total == 0 || ns == 0 || nd_real == 0 || nd_cells == 0;
    if (job->empty) {
        rc = nnps_clear(ctx, nd_total, counts, slice_off);  // every row empty
        if (rc) return rc;
        if (dev) {
            hipLaunchKernelGGL(k_nnps_dev_totals, dim3(1), dim3(64), 0, ctx->stream,
                               (const int64_t *)nullptr, (const int64_t *)nullptr,
                               nd_total == 0 ? (const int32_t *)nullptr : dev->ncells_dev, cap,
                               nd_cells, dev->totals, dev->status);
            SPH_LAUNCH_CHECK(ctx, "k_nnps_dev_totals");
        }
        return SPH_OK;
    }
    CellArgs &a = job->a;
    a = cell_args_base();
    a.g = make_griddev(ctx, grid);
    a.f32 = f32;
    a.drec = d_rec;
    a.drec32 = d_rec32;
    a.srec32 = s_rec32;
    a.ox = f32 ? ox : nullptr;
    a.oy = f32 ? oy : nullptr;
    a.oz = f32 ? oz : nullptr;
    a.dperm = d_perm;
    a.dstart = d_cell_start;
    a.dcells = d_cells;
    a.ncell_list = nd_cells;
    a.nd_real = nd_real;
    a.sstart = s_cell_start;
    a.srec = s_rec;
    a.rs = radius_scale;
    a.same = same_array;
    a.s_near = same_array ? nullptr : s_near;
    if (dev) {
        a.ncell_dev = dev->ncells_dev;
        a.cap = cap;
        a.status = dev->status;
        a.totals = dev->totals;  // the fill's first thread writes the totals and the status
        a.cell_bound = nd_cells;
    }
    return SPH_OK;
}

static int nnps_build(sph_ctx *ctx, const sph_grid *grid, int f32, const double *d_rec,
                      const float *d_rec32, const double *ox, const double *oy,
                      const double *oz, const int32_t *d_perm, const int32_t *d_cell_start,
                      const int32_t *d_cells, int64_t nd_cells, int64_t nd_total,
                      int64_t nd_real, const int32_t *s_cell_start, const double *s_rec,
                      const float *s_rec32, int64_t ns, double radius_scale, int same_array,
                      int32_t *counts, int64_t *slice_off, int32_t *indices, int64_t cap,
                      int64_t *slots_out, int64_t *nnz_out, const NnpsDevOut *dev,
                      const uint8_t *s_near = nullptr, int near_reach = 0)
{
    NnpsJob job;
    int rc = nnps_prepare(ctx, grid, f32, d_rec, d_rec32, ox, oy, oz, d_perm, d_cell_start,
                          d_cells, nd_cells, nd_total, nd_real, s_cell_start, s_rec, s_rec32, ns,
                          radius_scale, same_array, counts, slice_off, indices, cap, slots_out,
                          nnz_out, dev, s_near, near_reach, &job);
    if (rc || job.empty) return rc;
    CellArgs &a = job.a;
    const int64_t nslices = (nd_total + 63) / 64;
    int64_t *nnz_dev = nullptr;
    rc = nnps_count(ctx, a, nd_cells, nd_total, counts, slice_off, &nnz_dev);
    if (rc) return rc;
    if (dev) {
        a.nnz_dev = nnz_dev;
        return nnps_fill(ctx, a, nd_cells, nd_total, counts, slice_off, indices);
    }
    rc = sph_ensure_pinned(ctx, 2 * sizeof(int64_t));
    if (rc) return rc;
    SPH_HIP_CHECK(ctx, hipMemcpyAsync(ctx->pinned, slice_off + nslices, sizeof(int64_t),
                                      hipMemcpyDeviceToHost, ctx->stream));
    SPH_HIP_CHECK(ctx, hipMemcpyAsync((char *)ctx->pinned + sizeof(int64_t), nnz_dev,
                                      sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    SPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    const int64_t slots = ((int64_t *)ctx->pinned)[0];
    *slots_out = slots;
    *nnz_out = ((int64_t *)ctx->pinned)[1];
    if (slots > cap || !indices) {
        sph_set_error(ctx, SPH_ENEEDCAP, "neighbour capacity %lld < %lld", (long long)cap,
                      (long long)slots);
        return SPH_ENEEDCAP;
    }
    return nnps_fill(ctx, a, nd_cells, nd_total, counts, slice_off, indices);
}

static int nnps_prepare_dev(sph_ctx *ctx, const sph_grid *grid, const sph_nnps_dev_args *b,
                            const NnpsDevOut *dev, NnpsJob *job)
{
    SPH_REQUIRE(ctx, b, "sph_nnps_build_dev_pair: null build arguments");
    const int f32 = b->f32 != 0;
    return nnps_prepare(ctx, grid, f32, f32 ? nullptr : (const double *)b->d_rec,
                        f32 ? (const float *)b->d_rec : nullptr, b->d_x, b->d_y, b->d_z,
                        b->d_perm, b->d_cell_start, b->d_cells, b->max_cells, b->nd_total,
                        b->nd_real, b->s_cell_start, f32 ? nullptr : (const double *)b->s_rec,
                        f32 ? (const float *)b->s_rec : nullptr, b->ns, b->radius_scale,
                        b->same_array, b->counts, b->slice_off, b->indices, b->cap, nullptr,
                        nullptr, dev, b->s_near, b->near_reach, job);
}

// count_slices of two builds, each of its launches for both
static int pair_slices(sph_ctx *ctx, NnpsJob &j0, int64_t *sum0, size_t scan0, NnpsJob &j1,
                       int64_t *sum1, size_t scan1)
{
    const int64_t ns0 = j0.a.nslices, ns1 = j1.a.nslices;
    const int64_t nt0 = (ns0 + SLICE_TILE - 1) / SLICE_TILE;
    const int64_t nt1 = (ns1 + SLICE_TILE - 1) / SLICE_TILE;
    if (ns0 == 0 || ns1 == 0 || nt0 > SLICE_MAX_TILES || nt1 > SLICE_MAX_TILES) {
        int rc = count_slices(ctx, j0.a, j0.counts, j0.slice_off, sum0, scan0, j0.nd_total);
        if (rc) return rc;
        return count_slices(ctx, j1.a, j1.counts, j1.slice_off, sum1, scan1, j1.nd_total);
    }
    const SliceArgs s0 = {j0.counts, j0.nd_total, ns0, j0.slice_off, sum0, sum0 + ns0 + 1};
    const SliceArgs s1 = {j1.counts, j1.nd_total, ns1, j1.slice_off, sum1, sum1 + ns1 + 1};
    const int64_t nsl = ns0 > ns1 ? ns0 : ns1;
    hipLaunchKernelGGL(k_slices_pair, dim3(sph_blocks(nsl * 64, NN_THREADS), 2),
                       dim3(NN_THREADS), 0, ctx->stream, s0, s1);
    SPH_LAUNCH_CHECK(ctx, "k_slices_pair");
    // the tile sums after the copy of the widths (16-byte aligned: the sums are)
    const SliceScan t0 = {sum0 + ns0 + 1, sum0, ns0, nt0, (longlong2 *)(sum0 + 2 * ns0 + 2),
                          j0.slice_off, sum0 + ns0};
    const SliceScan t1 = {sum1 + ns1 + 1, sum1, ns1, nt1, (longlong2 *)(sum1 + 2 * ns1 + 2),
                          j1.slice_off, sum1 + ns1};
    const unsigned ntl = (unsigned)(nt0 > nt1 ? nt0 : nt1);
    hipLaunchKernelGGL(k_slice_tile_sums_pair, dim3(ntl, 2), dim3(SLICE_THREADS), 0, ctx->stream,
                       t0, t1);
    SPH_LAUNCH_CHECK(ctx, "k_slice_tile_sums_pair");
    hipLaunchKernelGGL(k_slice_tile_scan_pair, dim3(ntl, 2), dim3(SLICE_THREADS), 0, ctx->stream,
                       t0, t1);
    SPH_LAUNCH_CHECK(ctx, "k_slice_tile_scan_pair");
    return SPH_OK;
}

// blocks of the first build of a paired launch: a multiple of the XCD count,
// so the second build's block b - g0 runs on the XCD block b does
static unsigned pair_split(int64_t g)
{
    return (unsigned)((g + SPH_NXCD - 1) / SPH_NXCD * SPH_NXCD);
}

extern "C" int sph_nnps_build_dev_pair(sph_ctx *ctx, const sph_grid *grid,
                                       const sph_nnps_dev_args *b0,
                                       const sph_nnps_dev_args *b1)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, b0 && b1, "sph_nnps_build_dev_pair: null build arguments");
    SPH_REQUIRE(ctx, (b0->f32 != 0) == (b1->f32 != 0),
                "sph_nnps_build_dev_pair: both builds must have one precision");
    SPH_REQUIRE(ctx, b0->counts != b1->counts && b0->slice_off != b1->slice_off &&
                         b0->indices != b1->indices && b0->totals != b1->totals,
                "sph_nnps_build_dev_pair: the two builds share an output buffer");
    const NnpsDevOut dev0 = {b0->nd_cells_dev, b0->totals, b0->status};
    const NnpsDevOut dev1 = {b1->nd_cells_dev, b1->totals, b1->status};
    NnpsJob j0, j1;
    int rc = nnps_prepare_dev(ctx, grid, b0, &dev0, &j0);
    if (rc) return rc;
    rc = nnps_prepare_dev(ctx, grid, b1, &dev1, &j1);
    if (rc) return rc;
    if (j0.empty || j1.empty) {  // one build left (or none): its own launches
        NnpsJob *j = j0.empty ? &j1 : &j0;
        if (j->empty) return SPH_OK;
        int64_t *nnz_dev = nullptr;
        rc = nnps_count(ctx, j->a, j->nd_cells, j->nd_total, j->counts, j->slice_off, &nnz_dev);
        if (rc) return rc;
        j->a.nnz_dev = nnz_dev;
        return nnps_fill(ctx, j->a, j->nd_cells, j->nd_total, j->counts, j->slice_off,
                         j->indices);
    }
    // one scratch allocation for both, so neither carve moves
    const CountScratch c0 = count_scratch(j0.a, j0.nd_cells, j0.nd_total);
    const CountScratch c1 = count_scratch(j1.a, j1.nd_cells, j1.nd_total);
    rc = sph_ensure_scratch(ctx, c0.total + c1.total);
    if (rc) return rc;
    int64_t *sum0, *sum1;
    size_t scan0, scan1;
    count_carve(ctx, j0.a, c0, 0, j0.nd_total, &sum0, &scan0);
    count_carve(ctx, j1.a, c1, c0.total, j1.nd_total, &sum1, &scan1);
    j0.a.nnz_dev = sum0 + j0.a.nslices;
    j1.a.nnz_dev = sum1 + j1.a.nslices;
    const int64_t nd_max = j0.nd_total > j1.nd_total ? j0.nd_total : j1.nd_total;
    hipLaunchKernelGGL(k_nnps_reset_pair, dim3(sph_blocks(nd_max, NN_THREADS), 2),
                       dim3(NN_THREADS), 0, ctx->stream, j0.counts, j0.nd_total, j0.a.novf,
                       j1.counts, j1.nd_total, j1.a.novf);
    SPH_LAUNCH_CHECK(ctx, "k_nnps_reset_pair");
    const unsigned g0 = pair_split(count_grid(j0.a, j0.nd_cells));
    const int64_t g1 = count_grid(j1.a, j1.nd_cells);
    auto count_kern = j0.a.f32 ? k_nnps_cells_pair<float> : k_nnps_cells_pair<double>;
    hipLaunchKernelGGL(count_kern, dim3((unsigned)(g0 + g1)), dim3(CW), 0, ctx->stream, j0.a,
                       j0.counts, j1.a, j1.counts, g0);
    SPH_LAUNCH_CHECK(ctx, "k_nnps_cells_pair<count>");
    rc = pair_slices(ctx, j0, sum0, scan0, j1, sum1, scan1);
    if (rc) return rc;
    const unsigned f0 = pair_split(sph_blocks(j0.nd_total, NN_THREADS));
    const int64_t f1 = sph_blocks(j1.nd_total, NN_THREADS);
    hipLaunchKernelGGL(k_nnps_fill_pair, dim3((unsigned)(f0 + f1)), dim3(NN_THREADS), 0,
                       ctx->stream, j0.a, j0.nd_total, j0.counts, (const int64_t *)j0.slice_off,
                       j0.indices, j1.a, j1.nd_total, j1.counts, (const int64_t *)j1.slice_off,
                       j1.indices, f0);
    SPH_LAUNCH_CHECK(ctx, "k_nnps_fill_pair");
    // the cells the count pass flagged
    const int64_t ncl = j0.nd_cells > j1.nd_cells ? j0.nd_cells : j1.nd_cells;
    const int64_t flag_grid = ncl < NN_OVF_GRID ? ncl : NN_OVF_GRID;
    auto flag_kern = j0.a.f32 ? k_nnps_flagged_pair<float> : k_nnps_flagged_pair<double>;
    hipLaunchKernelGGL(flag_kern, dim3((unsigned)flag_grid, 2), dim3(CW), 0, ctx->stream, j0.a,
                       j0.counts, (const int64_t *)j0.slice_off, j0.indices, j1.a, j1.counts,
                       (const int64_t *)j1.slice_off, j1.indices);
    SPH_LAUNCH_CHECK(ctx, "k_nnps_flagged_pair");
    return SPH_OK;
}

extern "C" int sph_nnps_build(sph_ctx *ctx, const sph_grid *grid, const double *d_rec,
                              const int32_t *d_perm, const int32_t *d_cell_start,
                              const int32_t *d_cells, int64_t nd_cells, int64_t nd_total,
                              int64_t nd_real, const int32_t *s_cell_start, const double *s_rec,
                              int64_t ns, double radius_scale, int same_array, int32_t *counts,
                              int64_t *slice_off, int32_t *indices, int64_t cap,
                              int64_t *slots_out, int64_t *nnz_out)
{
    return nnps_build(ctx, grid, 0, d_rec, nullptr, nullptr, nullptr, nullptr, d_perm, d_cell_start,
                      d_cells, nd_cells, nd_total, nd_real, s_cell_start, s_rec, nullptr, ns,
                      radius_scale, same_array, counts, slice_off, indices, cap, slots_out,
                      nnz_out, nullptr);
}

extern "C" int sph_nnps_build_dev(sph_ctx *ctx, const sph_grid *grid, const double *d_rec,
                                  const int32_t *d_perm, const int32_t *d_cell_start,
                                  const int32_t *d_cells, const int32_t *nd_cells_dev,
                                  int64_t max_cells, int64_t nd_total, int64_t nd_real,
                                  const int32_t *s_cell_start, const double *s_rec, int64_t ns,
                                  double radius_scale, int same_array, int32_t *counts,
                                  int64_t *slice_off, int32_t *indices, int64_t cap,
                                  int64_t *totals, int32_t *status, const uint8_t *s_near,
                                  int near_reach)
{
    const NnpsDevOut dev = {nd_cells_dev, totals, status};
    return nnps_build(ctx, grid, 0, d_rec, nullptr, nullptr, nullptr, nullptr, d_perm,
                      d_cell_start, d_cells, max_cells, nd_total, nd_real, s_cell_start, s_rec,
                      nullptr, ns, radius_scale, same_array, counts, slice_off, indices, cap,
                      nullptr, nullptr, &dev, s_near, near_reach);
}

extern "C" int sph_nnps_build_f32(sph_ctx *ctx, const sph_grid *grid, const float *d_rec,
                                  const double *d_x, const double *d_y, const double *d_z,
                                  const int32_t *d_perm, const int32_t *d_cell_start,
                                  const int32_t *d_cells, int64_t nd_cells, int64_t nd_total,
                                  int64_t nd_real, const int32_t *s_cell_start,
                                  const float *s_rec, int64_t ns, double radius_scale,
                                  int same_array, int32_t *counts, int64_t *slice_off,
                                  int32_t *indices, int64_t cap, int64_t *slots_out,
                                  int64_t *nnz_out)
{
    return nnps_build(ctx, grid, 1, nullptr, d_rec, d_x, d_y, d_z,
                      d_perm, d_cell_start, d_cells, nd_cells, nd_total, nd_real, s_cell_start,
                      nullptr, s_rec, ns, radius_scale, same_array, counts, slice_off, indices,
                      cap, slots_out, nnz_out, nullptr);
}

extern "C" int sph_nnps_build_f32_dev(sph_ctx *ctx, const sph_grid *grid, const float *d_rec,
                                      const double *d_x, const double *d_y, const double *d_z,
                                      const int32_t *d_perm, const int32_t *d_cell_start,
                                      const int32_t *d_cells, const int32_t *nd_cells_dev,
                                      int64_t max_cells, int64_t nd_total, int64_t nd_real,
                                      const int32_t *s_cell_start, const float *s_rec,
                                      int64_t ns, double radius_scale, int same_array,
                                      int32_t *counts, int64_t *slice_off, int32_t *indices,
                                      int64_t cap, int64_t *totals, int32_t *status,
                                      const uint8_t *s_near, int near_reach)
{
    const NnpsDevOut dev = {nd_cells_dev, totals, status};
    return nnps_build(ctx, grid, 1, nullptr, d_rec, d_x, d_y, d_z, d_perm, d_cell_start, d_cells,
                      max_cells, nd_total, nd_real, s_cell_start, nullptr, s_rec, ns,
                      radius_scale, same_array, counts, slice_off, indices, cap, nullptr,
                      nullptr, &dev, s_near, near_reach);
}

// ---------------------------------------------------------------------------
// cells within reach of a source array
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(NN_THREADS)
    k_fill_u8(uint8_t *__restrict__ p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * NN_THREADS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * NN_THREADS)
        p[i] = 0;
}

// one lane per non-empty source cell: the cell's coordinates from its first
// particle (the double coordinates the binning used), then near = 1 for
// every cell of the box of half-width `reach` around it
__global__ void __launch_bounds__(NN_THREADS)
    k_cell_near(GridDev g, const int32_t *__restrict__ cells, const int32_t *__restrict__ ncells_dev,
                int64_t ncells_max, const int32_t *__restrict__ cell_start,
                const int32_t *__restrict__ perm, const double *__restrict__ x,
                const double *__restrict__ y, const double *__restrict__ z, int reach,
                uint8_t *__restrict__ near)
{
    int64_t nc = ncells_max;
    if (ncells_dev && (int64_t)*ncells_dev < nc) nc = *ncells_dev;
    const int64_t w = (int64_t)blockIdx.x * NN_THREADS + threadIdx.x;
    if (w >= nc) return;
    const int32_t i = perm[cell_start[cells[w]]];
    const int cx = min(max(cell_coord(x[i], g.cs) - g.cmin0, 0), g.d0 - 1);
    const int cy = min(max(cell_coord(y[i], g.cs) - g.cmin1, 0), g.d1 - 1);
    const int cz = min(max(cell_coord(z[i], g.cs) - g.cmin2, 0), g.d2 - 1);
    for (int ix = max(cx - reach, 0); ix <= min(cx + reach, g.d0 - 1); ++ix)
        for (int iy = max(cy - reach, 0); iy <= min(cy + reach, g.d1 - 1); ++iy)
            for (int iz = max(cz - reach, 0); iz <= min(cz + reach, g.d2 - 1); ++iz)
                near[cell_key(g, ix, iy, iz)] = 1;
}

extern "C" int sph_cell_near(sph_ctx *ctx, const sph_grid *grid, const int32_t *s_cell_start,
                             const int32_t *s_cells, const int32_t *ns_cells_dev,
                             int64_t ns_cells, const int32_t *s_perm, const double *s_x,
                             const double *s_y, const double *s_z, int reach, uint8_t *near)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, grid && near && reach >= 0 && reach <= 8 && ns_cells >= 0,
                "sph_cell_near: bad argument");
    SPH_REQUIRE(ctx, ns_cells == 0 || (s_cell_start && s_cells && s_perm && s_x && s_y && s_z),
                "sph_cell_near: null source arrays");
    int rc = sph_upload_key_tables(ctx, grid);
    if (rc) return rc;
    const GridDev g = make_griddev(ctx, grid);
    const int64_t nk = grid->ncells;
    int64_t fb = (nk + NN_THREADS - 1) / NN_THREADS;
    if (fb > 4096) fb = 4096;
    hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)(fb > 0 ? fb : 1)), dim3(NN_THREADS), 0,
                       ctx->stream, near, nk);
    SPH_LAUNCH_CHECK(ctx, "k_fill_u8");
    if (ns_cells == 0) return SPH_OK;
    hipLaunchKernelGGL(k_cell_near, dim3(sph_blocks(ns_cells, NN_THREADS)), dim3(NN_THREADS), 0,
                       ctx->stream, g, s_cells, ns_cells_dev, ns_cells, s_cell_start, s_perm,
                       s_x, s_y, s_z, reach, near);
    SPH_LAUNCH_CHECK(ctx, "k_cell_near");
    return SPH_OK;
}

int sph_nnps_build_batch(sph_ctx *ctx, const GridDev *grids, const int32_t *sim_of,
                         const int32_t *key_sim,
                         const double *rec, const float *rec32, const double *ox,
                         const double *oy, const double *oz, const int32_t *perm,
                         const int32_t *cell_start,
                         const int32_t *cells, const int32_t *ncell_dev, int64_t max_cells,
                         int64_t n, double radius_scale, int maxch, int32_t *counts,
                         int64_t *slice_off, int32_t *indices, int64_t cap, int32_t *status,
                         int64_t *nnz_acc, int32_t trash, int cells_per_block)
{
    SPH_REQUIRE(ctx, maxch >= 1 && maxch <= NN_MAXCH, "sph_nnps_build_batch: bad maxch");
    SPH_REQUIRE(ctx, cells_per_block >= 1, "sph_nnps_build_batch: bad cells_per_block");
    int rc;
    if (n == 0 || max_cells == 0) return nnps_clear(ctx, n, counts, slice_off);
    CellArgs a = cell_args_base();
    a.grids = grids;
    a.sim_of = sim_of;
    a.key_sim = key_sim;
    a.ncell_dev = ncell_dev;
    a.f32 = rec32 != nullptr;
    a.drec = rec;
    a.drec32 = rec32;
    a.srec32 = rec32;
    if (a.f32) {
        a.ox = ox;
        a.oy = oy;
        a.oz = oz;
    }
    a.dperm = perm;
    a.dstart = cell_start;
    a.dcells = cells;
    a.ncell_list = max_cells;
    a.nd_real = n;
    a.sstart = cell_start;
    a.srec = rec;
    a.rs = radius_scale;
    a.same = 1;
    a.maxch = maxch;
    a.cap = cap;
    a.status = status;
    a.trash = trash;
    a.cells_per_block = cells_per_block;
    int64_t *nnz_dev = nullptr;
    rc = nnps_count(ctx, a, max_cells, n, counts, slice_off, &nnz_dev);
    if (rc) return rc;
    hipLaunchKernelGGL(k_nnps_dev_status, dim3(1), dim3(64), 0, ctx->stream,
                       (const int64_t *)(slice_off + a.nslices), (const int64_t *)nnz_dev, cap,
                       status, nnz_acc);
    SPH_LAUNCH_CHECK(ctx, "k_nnps_dev_status");
    return nnps_fill(ctx, a, max_cells, n, counts, slice_off, indices);
}

// ---------------------------------------------------------------------------
// export to original index space
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(NN_THREADS)
    k_export_counts(int64_t nd_real, const int32_t *__restrict__ rank,
                    const int32_t *__restrict__ counts, int64_t *__restrict__ out_counts)
{
    int64_t i = (int64_t)blockIdx.x * NN_THREADS + threadIdx.x;
    if (i >= nd_real) return;
    out_counts[i] = counts[rank[i]];
}

__global__ void __launch_bounds__(NN_THREADS)
    k_export_rows(int64_t nd_real, const int32_t *__restrict__ rank,
                  const int32_t *__restrict__ s_perm, const int32_t *__restrict__ counts,
                  const int64_t *__restrict__ slice_off, const int32_t *__restrict__ indices,
                  const int64_t *__restrict__ out_offsets, int32_t *__restrict__ out_indices)
{
    int64_t i = (int64_t)blockIdx.x * NN_THREADS + threadIdx.x;
    if (i >= nd_real) return;
    const int32_t t = rank[i];
    const int32_t c = counts[t];
    const int32_t *src = indices + slice_off[t >> 6] + (t & 63);
    int32_t *dst = out_indices + out_offsets[i];
    for (int32_t k = 0; k < c; ++k) dst[k] = s_perm[src[(int64_t)k << 6]];
}

extern "C" int sph_nnps_export(sph_ctx *ctx, int64_t nd_real, const int32_t *d_rank,
                               const int32_t *s_perm, const int32_t *counts,
                               const int64_t *slice_off, const int32_t *indices,
                               int64_t *out_offsets, int32_t *out_indices)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, nd_real >= 0 && out_offsets, "sph_nnps_export: bad arguments");
    if (nd_real == 0) {
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(NN_THREADS), 0, ctx->stream,
                           (int32_t *)out_offsets, 2, 0);
        SPH_LAUNCH_CHECK(ctx, "k_fill_i32");
        return SPH_OK;
    }
    SPH_REQUIRE(ctx, d_rank && s_perm && counts && slice_off && indices && out_indices,
                "sph_nnps_export: null array");
    hipLaunchKernelGGL(k_export_counts, dim3(sph_blocks(nd_real, NN_THREADS)), dim3(NN_THREADS),
                       0, ctx->stream, nd_real, d_rank, counts, out_offsets);
    SPH_LAUNCH_CHECK(ctx, "k_export_counts");
    int rc = sph_scan_exclusive_i64(ctx, out_offsets, nd_real);
    if (rc) return rc;
    hipLaunchKernelGGL(k_export_rows, dim3(sph_blocks(nd_real, NN_THREADS)), dim3(NN_THREADS), 0,
                       ctx->stream, nd_real, d_rank, s_perm, counts, slice_off, indices,
                       out_offsets, out_indices);
    SPH_LAUNCH_CHECK(ctx, "k_export_rows");
    return SPH_OK;
}

// ---------------------------------------------------------------------------
// symmetric variable-h lists (VarHNbrParticleLocator, base/nnps.pyx:691-786)
// ---------------------------------------------------------------------------
// Row of destination d = its own query (sources within h_d*rs, stencil order,
// own index not moved) followed by the real sources j outside that radius
// whose own radius h_j*rs reaches d, in ascending original index: the rows
// the reference builds by appending each source of the reversed locator's
// rows in ascending source order.  One lane per sorted destination walks the
// stencil of the largest reach (ceil(max_h*rs/cell), max_h bounding every h
// of both arrays); a destination's own stencil is a sub-box of it, visited
// in the same lexicographic order.
struct VarhArgs {
    GridDev g;
    const double *drec, *srec;
    // float records (single precision): the distance tests run in float and
    // the destination's cell comes from its double coordinates ox, oy, oz
    // (original order), the ones the binning used
    const float *drec32, *srec32;
    const double *ox, *oy, *oz;
    const int32_t *dperm, *sperm, *srank, *sstart;
    int64_t nd_total, nd_real, ns_real;
    double rs, rmax2;  // rmax2 = (max_h*rs)^2 bounds every source radius
    int reach;
    int32_t *fcount;   // per destination: entries of its own query
};

// appended entries a fill lane sorts in LDS (one column per lane, 40 KB per
// block); a longer appended part is sorted in place in the row
#define VARH_LDS 40

template <typename T> __device__ __forceinline__ const T *varh_drec(const VarhArgs &a);
template <> __device__ __forceinline__ const double *varh_drec<double>(const VarhArgs &a)
{
    return a.drec;
}
template <> __device__ __forceinline__ const float *varh_drec<float>(const VarhArgs &a)
{
    return a.drec32;
}
template <typename T> __device__ __forceinline__ const T *varh_srec(const VarhArgs &a);
template <> __device__ __forceinline__ const double *varh_srec<double>(const VarhArgs &a)
{
    return a.srec;
}
template <> __device__ __forceinline__ const float *varh_srec<float>(const VarhArgs &a)
{
    return a.srec32;
}

// dcell_of, cell_ovf (both or neither): only the destinations of the cells
// the per-cell pass flagged (cell_ovf < 0) are handled; the other rows are
// that pass's.
template <bool FILL, typename T>
__global__ void __launch_bounds__(NN_THREADS)
    k_varh_rows(VarhArgs a, int32_t *__restrict__ counts, const int64_t *__restrict__ slice_off,
                int32_t *__restrict__ indices, const int32_t *__restrict__ dcell_of,
                const int32_t *__restrict__ cell_ovf)
{
    __shared__ int32_t s_app[FILL ? VARH_LDS : 1][NN_THREADS];
    const int64_t t = (int64_t)sph_xcd_tile(blockIdx.x, gridDim.x) * NN_THREADS + threadIdx.x;
    if (t >= a.nd_total) return;
    if (dcell_of) {
        const int32_t w = dcell_of[t];
        if (w < 0 || cell_ovf[w] >= 0) return;
    }
    if (a.dperm[t] >= a.nd_real) {  // non-LocalReal destination: empty row
        if (!FILL) {
            counts[t] = 0;
            a.fcount[t] = 0;
        }
        return;
    }
    const T *dr = varh_drec<T>(a) + t * SPH_REC;
    const T *srec = varh_srec<T>(a);
    const T px = dr[SPH_R_X], py = dr[SPH_R_Y], pz = dr[SPH_R_Z];
    const T radius = dr[SPH_R_H] * T(a.rs);
    const T r2 = radius * radius;
    const T rmax2 = T(a.rmax2);
    const GridDev &g = a.g;
    double qx = px, qy = py, qz = pz;
    if (a.ox) {
        const int32_t i0 = a.dperm[t];
        qx = a.ox[i0];
        qy = a.oy[i0];
        qz = a.oz[i0];
    }
    const int cx = min(max(cell_coord(qx, g.cs) - g.cmin0, 0), g.d0 - 1);
    const int cy = min(max(cell_coord(qy, g.cs) - g.cmin1, 0), g.d1 - 1);
    const int cz = min(max(cell_coord(qz, g.cs) - g.cmin2, 0), g.d2 - 1);
    const int R = a.reach;
    int32_t nf = 0, ne = 0;
    int32_t *row = nullptr;
    int32_t fbase = 0, n = 0;
    if (FILL) {
        n = counts[t];
        if (n == 0) return;
        row = indices + slice_off[t >> 6] + (t & 63);
        fbase = a.fcount[t];
    }
    // the count pass found n - fbase appended entries: every LDS index below
    // is < n - fbase <= VARH_LDS
    const bool in_lds = FILL && n - fbase <= VARH_LDS;
    for (int ix = max(cx - R, 0); ix <= min(cx + R, g.d0 - 1); ++ix)
        for (int iy = max(cy - R, 0); iy <= min(cy + R, g.d1 - 1); ++iy)
            for (int iz = max(cz - R, 0); iz <= min(cz + R, g.d2 - 1); ++iz) {
                const int32_t k = cell_key(g, ix, iy, iz);
                const int32_t s1 = a.sstart[k + 1];
                for (int32_t s = a.sstart[k]; s < s1; ++s) {
                    const T *sr = srec + (int64_t)s * SPH_REC;
                    // cPoint_distance2(src, dst), base/point.pxd:35-37
                    const T ex = sr[SPH_R_X] - px, ey = sr[SPH_R_Y] - py, ez = sr[SPH_R_Z] - pz;
                    const T d2 = ex * ex + ey * ey + ez * ez;
                    if (d2 < r2) {
                        // the fill repeats the count pass's tests, so its
                        // entries are the counted ones (the bounds only
                        // guard the row)
                        if (FILL && nf < fbase) row[(int64_t)nf << 6] = s;
                        ++nf;
                    } else if (d2 < rmax2) {
                        // the source's own radius: (h_j*rs)^2 <= rmax2, so
                        // no pair that reaches d is skipped by the bound; the
                        // source's original index is read for the hits only
                        const T rj = sr[SPH_R_H] * T(a.rs);
                        if (d2 < rj * rj) {
                            const int32_t j = a.sperm[s];
                            if (j < a.ns_real) {
                                if (FILL && fbase + ne < n) {  // original index
                                    if (in_lds)
                                        s_app[ne][threadIdx.x] = j;
                                    else
                                        row[(int64_t)(fbase + ne) << 6] = j;
                                }
                                ++ne;
                            }
                        }
                    }
                }
            }
    if (!FILL) {
        counts[t] = nf + ne;
        a.fcount[t] = nf;
        return;
    }
    // appended sources in ascending original index (insertion sort; the
    // appended part is short: it is empty where h does not vary), then
    // back to sorted positions
    ne = min(ne, n - fbase);
    if (in_lds) {
        for (int32_t i = 1; i < ne; ++i) {
            const int32_t v = s_app[i][threadIdx.x];
            int32_t q = i - 1;
            while (q >= 0 && s_app[q][threadIdx.x] > v) {
                s_app[q + 1][threadIdx.x] = s_app[q][threadIdx.x];
                --q;
            }
            s_app[q + 1][threadIdx.x] = v;
        }
        for (int32_t i = 0; i < ne; ++i)
            row[(int64_t)(fbase + i) << 6] = a.srank[s_app[i][threadIdx.x]];
        return;
    }
    for (int32_t i = 1; i < ne; ++i) {
        const int32_t v = row[(int64_t)(fbase + i) << 6];
        int32_t q = i - 1;
        while (q >= 0 && row[(int64_t)(fbase + q) << 6] > v) {
            row[(int64_t)(fbase + q + 1) << 6] = row[(int64_t)(fbase + q) << 6];
            --q;
        }
        row[(int64_t)(fbase + q + 1) << 6] = v;
    }
    for (int32_t i = 0; i < ne; ++i) {
        int32_t *e = row + ((int64_t)(fbase + i) << 6);
        *e = a.srank[*e];
    }
}

// ---------------------------------------------------------------------------
// symmetric variable-h lists, one wavefront per destination cell
// ---------------------------------------------------------------------------
// The count pass of the fixed-h lists with two hit masks per (destination,
// 64-candidate chunk): own (d^2 < (h_d*rs)^2: the row's first part, in
// stencil order) and appended (outside that radius, inside the source's own
// radius h_j*rs, real sources only).  Candidates are the stencil box of the
// largest reach around the cell, visited in the lexicographic order of the
// box, which contains every destination's own stencil.  A cell whose box has
// more than MAXSEG cells or more than maxch chunks is flagged (cell_ovf = -1)
// and k_varh_rows builds its destinations' rows.  The distance and radius
// expressions are k_varh_rows's, so both give the same rows.
template <typename T>
__device__ void varh_cell(const CellArgs &a, int64_t w, int32_t *__restrict__ counts,
                          int32_t *seg_start, int32_t *seg_end, typename Quad<T>::type *dsts)
{
    const T *drec = drec_of<T>(a);
    const T *srec = srec_of<T>(a);
    const int lane = threadIdx.x;
    const int32_t key = a.dcells[w];
    const int32_t ds = a.dstart[key], de = a.dstart[key + 1];
    const GridDev &g = a.g;
    // the cell's coordinates, recomputed exactly as the binning did
    double c0x, c0y, c0z;
    if (a.ox) {
        const int32_t i0 = a.dperm[ds];
        c0x = a.ox[i0];
        c0y = a.oy[i0];
        c0z = a.oz[i0];
    } else {
        const T *r0 = drec + (int64_t)ds * SPH_REC;
        c0x = r0[SPH_R_X];
        c0y = r0[SPH_R_Y];
        c0z = r0[SPH_R_Z];
    }
    const int cx = min(max(cell_coord(c0x, g.cs) - g.cmin0, 0), g.d0 - 1);
    const int cy = min(max(cell_coord(c0y, g.cs) - g.cmin1, 0), g.d1 - 1);
    const int cz = min(max(cell_coord(c0z, g.cs) - g.cmin2, 0), g.d2 - 1);
    const int R = a.reach;
    const int x0 = max(cx - R, 0), x1 = min(cx + R, g.d0 - 1);
    const int y0 = max(cy - R, 0), y1 = min(cy + R, g.d1 - 1);
    const int z0 = max(cz - R, 0), z1 = min(cz + R, g.d2 - 1);
    const int nx = x1 - x0 + 1, ny = y1 - y0 + 1, nz = z1 - z0 + 1;
    const int nst = nx * ny * nz;
    int32_t st = 0, cnt = 0;
    if (nst <= MAXSEG && lane < nst) {
        const int iz = lane % nz, iy = (lane / nz) % ny, ix = lane / (nz * ny);
        const int32_t k = cell_key(g, x0 + ix, y0 + iy, z0 + iz);
        st = a.sstart[k];
        cnt = a.sstart[k + 1] - st;
    }
    const int incl = wave_incl_scan_i(cnt);
    const int total = __shfl(incl, CW - 1, 64);
    const int nch = (total + CW - 1) / CW;
    if (nst > MAXSEG || nch > a.maxch) {  // the lane-per-destination rows
        for (int32_t t = ds + lane; t < de; t += CW) a.dcell_of[t] = (int32_t)w;
        if (lane == 0) a.cell_ovf[w] = -1;
        return;
    }
    __syncthreads();  // the previous cell's segments are no longer read
    seg_start[lane] = st;  // MAXSEG == CW: one segment per lane
    seg_end[lane] = incl;
    __syncthreads();
    for (int32_t g0 = ds; g0 < de; g0 += CW) {  // destination groups of <= 64
        const int gn = min(CW, de - g0);
        const int32_t my_t = g0 + lane;
        bool mreal = false;
        if (lane < gn) {
            const T *dr = drec + (int64_t)my_t * SPH_REC;
            const T radius = dr[SPH_R_H] * T(a.rs);
            dsts[lane] = Quad<T>::make(dr[SPH_R_X], dr[SPH_R_Y], dr[SPH_R_Z], radius * radius);
            mreal = a.dperm[my_t] < a.nd_real;
        }
        __syncthreads();
        const uint64_t realmask = __ballot(mreal);
        int32_t nf = 0, ne = 0;
        for (int c = 0; c < nch; ++c) {
            const int q = c * CW + lane;
            const bool valid = q < total;
            int32_t p = 0;
            T sx = T(0.0), sy = T(0.0), sz = T(0.0), rj2 = T(0.0);
            bool jreal = false;
            if (valid) {
                int lo = 0, hi = MAXSEG - 1;  // first segment with end > q
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (seg_end[mid] > q) hi = mid; else lo = mid + 1;
                }
                const int prev = lo > 0 ? seg_end[lo - 1] : 0;
                p = seg_start[lo] + (q - prev);
                const T *sr = srec + (int64_t)p * SPH_REC;
                sx = sr[SPH_R_X];
                sy = sr[SPH_R_Y];
                sz = sr[SPH_R_Z];
                const T rj = sr[SPH_R_H] * T(a.rs);
                rj2 = rj * rj;
                const int32_t j = a.sperm[p];
                jreal = j < a.ns_real;
                if (g0 == ds) {
                    a.cand[w * (a.maxch * CW) + q] = p;
                    a.candj[w * (a.maxch * CW) + q] = j;
                }
            }
            const uint64_t validm = __builtin_amdgcn_ballot_w64(valid);
            uint32_t olo = 0, ohi = 0, alo = 0, ahi = 0;
            for (int d = 0; d < gn; ++d) {
                if (!((realmask >> d) & 1)) continue;  // non-LocalReal: empty row
                const auto D = dsts[d];
                // cPoint_distance2(src, dst), base/point.pxd:35-37
                const T ex = sx - D.x, ey = sy - D.y, ez = sz - D.z;
                const T d2 = ex * ex + ey * ey + ez * ez;
                const bool own = d2 < D.w;
                const uint64_t om = __builtin_amdgcn_ballot_w64(own) & validm;
                const uint64_t am =
                    __builtin_amdgcn_ballot_w64(!own && jreal && d2 < rj2) & validm;
                wave_writelane2(olo, ohi, (uint32_t)om, (uint32_t)(om >> 32), d);
                wave_writelane2(alo, ahi, (uint32_t)am, (uint32_t)(am >> 32), d);
            }
            nf += __popc(olo) + __popc(ohi);
            ne += __popc(alo) + __popc(ahi);
            if (lane < gn) {
                a.masks[(int64_t)c * a.mstride + my_t] = ((uint64_t)ohi << 32) | olo;
                a.amasks[(int64_t)c * a.mstride + my_t] = ((uint64_t)ahi << 32) | alo;
            }
        }
        if (lane < gn) {
            counts[my_t] = nf + ne;
            a.fcount[my_t] = nf;
            a.dcell_of[my_t] = (int32_t)w;
        }
        __syncthreads();  // the next group's destinations overwrite dsts
    }
    if (lane == 0) a.cell_ovf[w] = nch;
}

template <typename T>
__global__ void __launch_bounds__(CW) k_varh_cells(CellArgs a, int32_t *__restrict__ counts)
{
    __shared__ int32_t seg_start[MAXSEG];
    __shared__ int32_t seg_end[MAXSEG];  // inclusive prefix of the segment sizes
    __shared__ typename Quad<T>::type dsts[CW];
    int64_t ncl = a.ncell_list;
    if (a.ncell_dev && (int64_t)*a.ncell_dev < ncl) ncl = *a.ncell_dev;
    // XCD x walks cells [x*ncl/8, (x+1)*ncl/8) with its blocks (k_nnps_cells)
    const unsigned G = gridDim.x;
    if (G < SPH_NXCD) {
        for (int64_t w = blockIdx.x; w < ncl; w += G)
            varh_cell<T>(a, w, counts, seg_start, seg_end, dsts);
        return;
    }
    const unsigned x = blockIdx.x % SPH_NXCD, k = blockIdx.x / SPH_NXCD;
    const unsigned per = (G - x + SPH_NXCD - 1) / SPH_NXCD;
    const int64_t lo = ncl * x / SPH_NXCD, hi = ncl * (x + 1) / SPH_NXCD;
    for (int64_t w = lo + k; w < hi; w += per)
        varh_cell<T>(a, w, counts, seg_start, seg_end, dsts);
}

// rows start empty, with no cell (destinations of cells past the launch bound)
__global__ void __launch_bounds__(NN_THREADS)
    k_varh_reset(int64_t nd_total, int32_t *__restrict__ counts, int32_t *__restrict__ fcount,
                 int32_t *__restrict__ dcell_of)
{
    const int64_t t = (int64_t)blockIdx.x * NN_THREADS + threadIdx.x;
    if (t >= nd_total) return;
    counts[t] = 0;
    fcount[t] = 0;
    dcell_of[t] = -1;
}

// Count pass per destination cell (k_varh_cells; the destinations of flagged
// cells by k_varh_rows), slice offsets and totals, the host check of the
// capacity, then the fill and the sort of long appended parts.  Scratch:
// [own masks | appended masks | cell flags | candidate sources | cell of
// destination | own-query entries | candidate original indices | long-row
// count and list | slice sums | scan partials].
template <typename T>
static int nnps_build_varh(sph_ctx *ctx, const sph_grid *grid, const T *d_rec, const double *d_x,
                           const double *d_y, const double *d_z, const int32_t *d_perm,
                           const int32_t *d_cell_start, const int32_t *d_cells,
                           const int32_t *d_ncells_dev, int64_t nd_cells, int64_t nd_total,
                           int64_t nd_real, const int32_t *s_cell_start, const T *s_rec,
                           const int32_t *s_perm, const int32_t *s_rank, int64_t ns,
                           int64_t ns_real, double radius_scale, double max_h, int32_t *counts,
                           int64_t *slice_off, int32_t *indices, int64_t cap,
                           int64_t *slots_out, int64_t *nnz_out)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, grid && counts && slice_off && slots_out && nnz_out,
                "sph_nnps_build_varh: null argument");
    SPH_REQUIRE(ctx, nd_total >= 0 && nd_total < ((int64_t)1 << 31) && nd_real >= 0 &&
                         nd_real <= nd_total && ns >= 0 && ns < ((int64_t)1 << 31) &&
                         ns_real >= 0 && ns_real <= ns && nd_cells >= 0 && nd_cells <= nd_total,
                "sph_nnps_build_varh: bad particle or cell counts");
    SPH_REQUIRE(ctx, ns == 0 || (s_cell_start && s_rec && s_perm && s_rank),
                "sph_nnps_build_varh: null source arrays");
    SPH_REQUIRE(ctx, nd_total == 0 || (d_rec && d_perm && d_cell_start && d_cells),
                "sph_nnps_build_varh: null destination arrays");
    const bool f32 = sizeof(T) == sizeof(float);
    SPH_REQUIRE(ctx, !f32 || nd_total == 0 || (d_x && d_y && d_z),
                "sph_nnps_build_varh_f32: null destination coordinates");
    *slots_out = 0;
    *nnz_out = 0;
    const int64_t nslices = (nd_total + 63) / 64;
    if (nd_total == 0 || ns == 0 || nd_real == 0 || nd_cells == 0)
        return nnps_clear(ctx, nd_total, counts, slice_off);
    SPH_REQUIRE(ctx, radius_scale >= 0.0 && grid->cell_size > 0.0 && max_h >= 0.0,
                "sph_nnps_build_varh: bad radius scale, max_h or grid");
    const double rmax = max_h * radius_scale;
    const double reach = ceil(rmax / grid->cell_size);
    SPH_REQUIRE(ctx, reach <= 256.0,
                "sph_nnps_build_varh: the largest radius spans %g cells (at most 256)", reach);
    int rc = sph_upload_key_tables(ctx, grid);
    if (rc) return rc;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const int maxch = NN_MAXCH;
    const size_t masks_bytes = al((size_t)maxch * nd_total * sizeof(uint64_t));
    const size_t ovf_bytes = al((size_t)nd_cells * sizeof(int32_t));
    const size_t cand_bytes = al((size_t)nd_cells * maxch * CW * sizeof(int32_t));
    const size_t dcell_bytes = al((size_t)nd_total * sizeof(int32_t));
    const size_t fc_bytes = al((size_t)nd_total * sizeof(int32_t));
    const size_t sums_bytes = slice_sums_bytes(nslices);
    const size_t long_bytes = al((size_t)(nd_total + 1) * sizeof(int32_t));
    const size_t sums_off =
        2 * masks_bytes + ovf_bytes + 2 * cand_bytes + dcell_bytes + fc_bytes + long_bytes;
    rc = sph_ensure_scratch(ctx, sums_off + sums_bytes +
                                     sph_scan_scratch_bytes(nslices, sizeof(int64_t)));
    if (rc) return rc;
    char *sc = (char *)ctx->scratch;
    CellArgs c = cell_args_base();
    c.g = make_griddev(ctx, grid);
    c.f32 = f32;
    if (f32) {
        c.drec32 = (const float *)d_rec;
        c.srec32 = (const float *)s_rec;
        c.ox = d_x;
        c.oy = d_y;
        c.oz = d_z;
    } else {
        c.drec = (const double *)d_rec;
        c.srec = (const double *)s_rec;
    }
    c.dperm = d_perm;
    c.dstart = d_cell_start;
    c.dcells = d_cells;
    c.ncell_list = nd_cells;
    c.ncell_dev = d_ncells_dev;
    c.nd_real = nd_real;
    c.sstart = s_cell_start;
    c.rs = radius_scale;
    c.sperm = s_perm;
    c.ns_real = ns_real;
    c.reach = (int)reach;
    c.maxch = maxch;
    c.masks = (uint64_t *)sc;
    c.amasks = (uint64_t *)(sc + masks_bytes);
    c.mstride = nd_total;
    c.cell_ovf = (int32_t *)(sc + 2 * masks_bytes);
    c.cand = (int32_t *)(sc + 2 * masks_bytes + ovf_bytes);
    c.dcell_of = (int32_t *)(sc + 2 * masks_bytes + ovf_bytes + cand_bytes);
    c.fcount = (int32_t *)(sc + 2 * masks_bytes + ovf_bytes + cand_bytes + dcell_bytes);
    c.candj = (int32_t *)(sc + 2 * masks_bytes + ovf_bytes + cand_bytes + dcell_bytes + fc_bytes);
    c.nlong = (int32_t *)(sc + 2 * masks_bytes + ovf_bytes + 2 * cand_bytes + dcell_bytes +
                          fc_bytes);
    c.long_rows = c.nlong + 1;
    int64_t *slice_sum = (int64_t *)(sc + sums_off);
    int64_t *nnz_dev = slice_sum + nslices;
    // the rows of flagged cells: the lane-per-destination kernel
    VarhArgs a;
    memset(&a, 0, sizeof(a));
    a.g = c.g;
    a.drec = c.drec;
    a.srec = c.srec;
    a.drec32 = c.drec32;
    a.srec32 = c.srec32;
    a.ox = c.ox;
    a.oy = c.oy;
    a.oz = c.oz;
    a.dperm = d_perm;
    a.sperm = s_perm;
    a.srank = s_rank;
    a.sstart = s_cell_start;
    a.nd_total = nd_total;
    a.nd_real = nd_real;
    a.ns_real = ns_real;
    a.rs = radius_scale;
    // float: the bound rounded up (it only prunes; the source's own radius
    // decides)
    a.rmax2 = f32 ? rmax * rmax * (1.0 + 0x1p-18) : rmax * rmax;
    a.reach = (int)reach;
    a.fcount = c.fcount;
    const unsigned dgrid = sph_blocks(nd_total, NN_THREADS);
    hipLaunchKernelGGL(k_varh_reset, dim3(dgrid), dim3(NN_THREADS), 0, ctx->stream, nd_total,
                       counts, c.fcount, c.dcell_of);
    SPH_LAUNCH_CHECK(ctx, "k_varh_reset");
    const int64_t cgrid = nd_cells < NN_CELL_GRID ? nd_cells : NN_CELL_GRID;
    hipLaunchKernelGGL(k_varh_cells<T>, dim3((unsigned)cgrid), dim3(CW), 0, ctx->stream, c,
                       counts);
    SPH_LAUNCH_CHECK(ctx, "k_varh_cells");
    auto count_rows = k_varh_rows<false, T>;
    hipLaunchKernelGGL(count_rows, dim3(dgrid), dim3(NN_THREADS), 0, ctx->stream, a, counts,
                       (const int64_t *)nullptr, (int32_t *)nullptr,
                       (const int32_t *)c.dcell_of, (const int32_t *)c.cell_ovf);
    SPH_LAUNCH_CHECK(ctx, "k_varh_rows<count>");
    hipLaunchKernelGGL(k_slices, dim3(sph_blocks(nslices * 64, NN_THREADS)), dim3(NN_THREADS), 0,
                       ctx->stream, counts, nd_total, nslices, slice_off, slice_sum,
                       slice_sum + nslices + 1);
    SPH_LAUNCH_CHECK(ctx, "k_slices");
    rc = slices_to_offsets(ctx, slice_off, slice_sum + nslices + 1, slice_sum, nslices, nnz_dev,
                           sums_off + sums_bytes);
    if (rc) return rc;
    rc = sph_ensure_pinned(ctx, 3 * sizeof(int64_t));
    if (rc) return rc;
    int64_t *hp = (int64_t *)ctx->pinned;
    SPH_HIP_CHECK(ctx, hipMemcpyAsync(hp, slice_off + nslices, sizeof(int64_t),
                                      hipMemcpyDeviceToHost, ctx->stream));
    SPH_HIP_CHECK(ctx, hipMemcpyAsync(hp + 1, nnz_dev, sizeof(int64_t), hipMemcpyDeviceToHost,
                                      ctx->stream));
    hp[2] = 0;
    if (d_ncells_dev)
        SPH_HIP_CHECK(ctx, hipMemcpyAsync(hp + 2, d_ncells_dev, sizeof(int32_t),
                                          hipMemcpyDeviceToHost, ctx->stream));
    SPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    const int64_t slots = hp[0];
    *slots_out = slots;
    *nnz_out = hp[1];
    const int64_t ncells_now = d_ncells_dev ? (int64_t)*(const int32_t *)(hp + 2) : nd_cells;
    if (ncells_now > nd_cells)
        return sph_set_error(ctx, SPH_EINVAL,
                             "sph_nnps_build_varh: %lld non-empty destination cells exceed the "
                             "bound %lld", (long long)ncells_now, (long long)nd_cells);
    if (slots > cap || !indices) {
        sph_set_error(ctx, SPH_ENEEDCAP, "neighbour capacity %lld < %lld", (long long)cap,
                      (long long)slots);
        return SPH_ENEEDCAP;
    }
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, ctx->stream, c.nlong, 1, 0);
    SPH_LAUNCH_CHECK(ctx, "k_fill_i32");
    hipLaunchKernelGGL(k_nnps_fill_slices<true>, dim3(dgrid), dim3(NN_THREADS), 0, ctx->stream,
                       c, nd_total, counts, (const int64_t *)slice_off, indices, s_rank);
    SPH_LAUNCH_CHECK(ctx, "k_nnps_fill_slices<varh>");
    hipLaunchKernelGGL(k_varh_long_rows, dim3(4096), dim3(CW), 0, ctx->stream, c,
                       (const int32_t *)counts, (const int64_t *)slice_off, indices, s_rank);
    SPH_LAUNCH_CHECK(ctx, "k_varh_long_rows");
    auto varh_fill_kern = k_varh_rows<true, T>;
    hipLaunchKernelGGL(varh_fill_kern, dim3(dgrid), dim3(NN_THREADS), 0, ctx->stream, a, counts,
                       (const int64_t *)slice_off, indices, (const int32_t *)c.dcell_of,
                       (const int32_t *)c.cell_ovf);
    SPH_LAUNCH_CHECK(ctx, "k_varh_rows<fill>");
    return SPH_OK;
}

extern "C" int sph_nnps_build_varh(sph_ctx *ctx, const sph_grid *grid, const double *d_rec,
                                   const int32_t *d_perm, const int32_t *d_cell_start,
                                   const int32_t *d_cells, const int32_t *nd_cells_dev,
                                   int64_t nd_cells, int64_t nd_total, int64_t nd_real,
                                   const int32_t *s_cell_start, const double *s_rec,
                                   const int32_t *s_perm, const int32_t *s_rank, int64_t ns,
                                   int64_t ns_real, double radius_scale, double max_h,
                                   int32_t *counts, int64_t *slice_off, int32_t *indices,
                                   int64_t cap, int64_t *slots_out, int64_t *nnz_out)
{
    return nnps_build_varh<double>(ctx, grid, d_rec, nullptr, nullptr, nullptr, d_perm,
                                   d_cell_start, d_cells, nd_cells_dev, nd_cells, nd_total,
                                   nd_real, s_cell_start, s_rec, s_perm, s_rank, ns, ns_real,
                                   radius_scale, max_h, counts, slice_off, indices, cap,
                                   slots_out, nnz_out);
}

extern "C" int sph_nnps_build_varh_f32(sph_ctx *ctx, const sph_grid *grid, const float *d_rec,
                                       const double *d_x, const double *d_y, const double *d_z,
                                       const int32_t *d_perm, const int32_t *d_cell_start,
                                       const int32_t *d_cells, const int32_t *nd_cells_dev,
                                       int64_t nd_cells, int64_t nd_total, int64_t nd_real,
                                       const int32_t *s_cell_start, const float *s_rec,
                                       const int32_t *s_perm, const int32_t *s_rank, int64_t ns,
                                       int64_t ns_real, double radius_scale, double max_h,
                                       int32_t *counts, int64_t *slice_off, int32_t *indices,
                                       int64_t cap, int64_t *slots_out, int64_t *nnz_out)
{
    return nnps_build_varh<float>(ctx, grid, d_rec, d_x, d_y, d_z, d_perm, d_cell_start, d_cells,
                                  nd_cells_dev, nd_cells, nd_total, nd_real, s_cell_start, s_rec,
                                  s_perm, s_rank, ns, ns_real, radius_scale, max_h, counts,
                                  slice_off, indices, cap, slots_out, nnz_out);
}

// The pencil walk of the LDS-staged neighbour loops (linear cell order),
// shared by the plain operators (sph_staged.hip) and the batched ones
// (sph_batch.hip).  See sph_staged.hip for the scheme.
#pragma once
#include "sph_internal.h"
#include "sph_terms.h"

#define STG_THREADS 64  // one wavefront per block: the staging needs no block barrier
#define STG_WIN 4       // list entries loaded ahead of the one being summed


__device__ __forceinline__ int stg_wave_max(int v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = max(v, __shfl_xor(v, d, 64));
    return v;
}

// inclusive prefix maximum over the lanes
__device__ __forceinline__ int32_t stg_prefix_max(int32_t v)
{
    const int lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int32_t u = __shfl_up(v, d, 64);
        if (lane >= d) v = max(v, u);
    }
    return v;
}

// stencil reach of a destination in cells, as the list count pass computes it
template <typename T> __device__ __forceinline__ int stg_reach(T h, double rs, double cs)
{
    double r = (double)h * rs / cs;
    if (sizeof(T) == sizeof(float)) r *= 1.0 - 0x1p-20;
    return r > 0.0 ? (r < 256.0 ? (int)ceil(r) : 256) : 0;
}

// The pencil pass of one source for one wavefront.  Per pencil (dx, dy) in
// stencil order: each lane's range [lo, hi) of sorted positions, the merged
// intervals staged into lds (CAP records of NP pairs, filled by `stage`), then
// the lane's entries inside its range summed by `term(j, pairs)` (pairs from
// LDS, or from global memory through `fetch`).  k counts the lane's consumed
// entries; jb is its window of the next entries.
template <typename T, int NP, int CAP, typename StageF, typename FetchF, typename TermF>
__device__ __forceinline__ void stg_row(const GridDev &g, int cx, int cy, int cz, int dmax,
                                        const int32_t *__restrict__ sstart,
                                        const int32_t *__restrict__ row, int32_t e,
                                        int32_t self_t, bool same, typename Pair<T>::type *lds,
                                        StageF stage, FetchF fetch, TermF term)
{
    using P = typename Pair<T>::type;
    const int lane = threadIdx.x;
    int32_t k = 0;
    int32_t jb[STG_WIN + 1];
#pragma unroll
    for (int u = 0; u <= STG_WIN; ++u)
        jb[u] = u < e ? __builtin_nontemporal_load(row + ((int64_t)u << 6)) : -1;
    // grids one cell deep in z (2D, 1D): the pencils run along y, one per dx
    // (wave-uniform: with a grid per lane, only when the grid of every lane
    // with entries is)
    const bool flat = !__builtin_amdgcn_ballot_w64(e > 0 && g.d2 != 1);
    const int dy0 = flat ? 0 : -dmax, dy1 = flat ? 0 : dmax;
    for (int dx = -dmax; dx <= dmax; ++dx) {
        for (int dy = dy0; dy <= dy1; ++dy) {
            const bool need = k < e;
            if (!__builtin_amdgcn_ballot_w64(need)) return;  // every row done
            int32_t lo = 0, hi = 0;
            const int px = cx + dx, py = cy + dy;
            if (need && px >= 0 && px < g.d0 && py >= 0 && py < g.d1) {
                if (flat) {
                    const int y0 = max(cy - dmax, 0), y1 = min(cy + dmax, g.d1 - 1);
                    lo = sstart[cell_key(g, px, y0, 0)];
                    hi = sstart[cell_key(g, px, y1, 0) + 1];
                } else {
                    const int z0 = max(cz - dmax, 0), z1 = min(cz + dmax, g.d2 - 1);
                    lo = sstart[cell_key(g, px, py, z0)];
                    hi = sstart[cell_key(g, px, py, z1) + 1];
                }
            }
            const bool has = hi > lo;
            if (!__builtin_amdgcn_ballot_w64(has)) continue;
            // merge the lanes' ranges (non-decreasing in lane order) into
            // intervals: a lane starts one when its lo lies past every earlier hi
            const int32_t pm = stg_prefix_max(has ? hi : -1);
            int32_t pmx = __shfl_up(pm, 1, 64);
            if (lane == 0) pmx = -1;
            const uint64_t starts = __builtin_amdgcn_ballot_w64(has && lo > pmx);
            int32_t total = 0;
            for (uint64_t r = starts; r;) {
                const int ls = __builtin_ctzll(r);
                r &= r - 1;
                const int le = r ? __builtin_ctzll(r) : 64;
                total += __builtin_amdgcn_readlane(pm, le - 1) -
                         __builtin_amdgcn_readlane(lo, ls);
            }
            const bool staged = total <= CAP;
            int32_t shift = 0;
            int32_t off = 0;
            for (uint64_t r = starts; r;) {
                const int ls = __builtin_ctzll(r);
                r &= r - 1;
                const int le = r ? __builtin_ctzll(r) : 64;
                const int32_t slo = __builtin_amdgcn_readlane(lo, ls);
                const int32_t len = __builtin_amdgcn_readlane(pm, le - 1) - slo;
                if (lane >= ls && lane < le) shift = off - slo;
                if (staged)
                    for (int32_t q = lane; q < len; q += STG_THREADS)
                        stage(lds + (int64_t)(off + q) * NP, slo + q);
                off += len;
            }
            __syncthreads();
            // this pencil's entries of every row, in list order
            bool run = need;
            while (__builtin_amdgcn_ballot_w64(run)) {
                if (run) {
                    const int32_t j = jb[0];
                    const bool inwin = j >= lo && j < hi;
                    bool take = inwin;
                    if (!inwin && same) {
                        if (k == e - 1)
                            take = j == self_t;  // the own index in the last slot
                        else  // the stencil-last entry swapped into the own index's slot
                            take = jb[1] < j && !(k + 1 == e - 1 && jb[1] == self_t);
                    }
                    if (take) {
                        P pr[NP];
                        if (inwin && staged) {
                            const P *src = lds + (int64_t)(j + shift) * NP;
#pragma unroll
                            for (int f = 0; f < NP; ++f) pr[f] = src[f];
                        } else {
                            fetch(pr, j);
                        }
                        term(pr);
                        ++k;
#pragma unroll
                        for (int u = 0; u < STG_WIN; ++u) jb[u] = jb[u + 1];
                        jb[STG_WIN] = k + STG_WIN < e
                                          ? __builtin_nontemporal_load(row + ((int64_t)(k + STG_WIN) << 6))
                                          : -1;
                        run = k < e;
                    } else {
                        run = false;
                    }
                }
            }
            __syncthreads();  // the next pencil's staging overwrites lds
        }
    }
    // what is left (the own index in the last slot after the last pencil with
    // staged ranges), still in list order, from global memory
    while (k < e) {
        P pr[NP];
        fetch(pr, jb[0]);
        term(pr);
        ++k;
#pragma unroll
        for (int u = 0; u < STG_WIN; ++u) jb[u] = jb[u + 1];
        jb[STG_WIN] =
            k + STG_WIN < e ? __builtin_nontemporal_load(row + ((int64_t)(k + STG_WIN) << 6)) : -1;
    }
}

// the destination's cell, recomputed as the binning computed it
__device__ __forceinline__ void stg_cell(const GridDev &g, double x, double y, double z, int &cx,
                                         int &cy, int &cz)
{
    cx = min(max(cell_coord(x, g.cs) - g.cmin0, 0), g.d0 - 1);
    cy = min(max(cell_coord(y, g.cs) - g.cmin1, 0), g.d1 - 1);
    cz = min(max(cell_coord(z, g.cs) - g.cmin2, 0), g.d2 - 1);
}


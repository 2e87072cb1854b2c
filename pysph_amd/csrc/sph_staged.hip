// SPHRho and MomentumEquation with the neighbour records staged in LDS
// (linear cell order).
//
// With cells keyed linearly, key = (cx*d1 + cy)*d2 + cz, the reference's row
// order -- stencil cells x -> y -> z, ascending index inside a cell
// (base/cell.pyx:95-118, base/nnps.pyx:271-356) -- is ascending sorted source
// position, apart from the own index swapped to the last slot
// (nnps.pyx:612-619).  The sources of one stencil pencil (cx+dx, cy+dy,
// cz-d..cz+d; on a grid one cell deep in z, cx+dx, cy-d..cy+d) are one
// contiguous range of sorted positions.  One wavefront
// takes the 64 destinations of a SELL-64 slice and walks the pencils in
// stencil order: per pencil it merges the 64 lanes' ranges (they are
// non-decreasing in lane order) into a few intervals, copies those records
// into LDS with contiguous loads, and each lane then sums the entries of its
// row that fall in its range, reading the record at LDS slot j + shift.  The
// lane stops at its first entry beyond the range and resumes at the next
// pencil, so every row is summed in list order with the arithmetic of
// sph_terms.h: the results equal k_summation_density's and k_momentum's bit
// for bit.  Out of range entries read from global memory instead: the
// swapped pair (the stencil-last entry in the own index's slot, recognised
// as the one place a same-array row decreases before its last entry, and the
// own index in the last slot) and every entry of a pencil whose intervals
// exceed the LDS capacity.
#include <string.h>

#include "sph_staged.h"

template <typename T> struct StagedSrc {
    SrcLists<T> L;
    const int32_t *sstart[SPH_MAX_SRC];  // cell starts of each source's binning
    int32_t same[SPH_MAX_SRC];           // source array == destination array
};

// ---------------------------------------------------------------------------
// SPHRho (density_funcs.pyx:37-73; the arithmetic of k_summation_density)
// ---------------------------------------------------------------------------
template <bool UH, typename T, int CAP>
__global__ void __launch_bounds__(STG_THREADS)
    k_density_staged(int dim, int64_t nd_total, int64_t nd_real, const int32_t *__restrict__ dperm,
                     StagedSrc<T> S, const T *__restrict__ drec, const double *__restrict__ ox,
                     const double *__restrict__ oy, const double *__restrict__ oz, GridDev g,
                     double rs, int accumulate, double *__restrict__ rho, T *__restrict__ out_s,
                     int64_t out_stride)
{
    using P = typename Pair<T>::type;
    constexpr int NP = UH ? 2 : 3;  // (x, y) (z, m) [(h, -)]
    __shared__ P lds[CAP * NP];
    const int lane = threadIdx.x;
    const int64_t slice = sph_xcd_tile(blockIdx.x, gridDim.x);
    const int64_t t = slice * STG_THREADS + lane;
    const bool act = t < nd_total;
    const int32_t i = act ? dperm[t] : 0;
    const bool real = act && i < nd_real;
    int32_t e[SPH_MAX_SRC];
    int32_t any = 0;
#pragma unroll
    for (int s = 0; s < SPH_MAX_SRC; ++s) {
        e[s] = (real && s < S.L.nsrc) ? S.L.counts[s][t] : 0;
        any |= e[s];
    }
    T val = T(0.0);
    if (real) val = accumulate ? (out_s ? out_s[t * out_stride] : T(rho[i])) : T(0.0);
    T xa = T(0), ya = T(0), za = T(0), ha = T(0), facA = T(0), invA = T(0);
    int cx = 0, cy = 0, cz = 0, dl = 0;
    if (any) {
        const auto *dr = rec2(drec, t);
        const auto dxy = dr[0], dzm = dr[1];
        xa = dxy.x;
        ya = dxy.y;
        za = dzm.x;
        ha = drec[t * SPH_REC + SPH_R_H];
        facA = sph_h_dim(ha, dim) * sph_cubic_fac1<T>(dim);
        invA = sph_inv_for_div(ha);
        if (ox) stg_cell(g, ox[i], oy[i], oz[i], cx, cy, cz);
        else stg_cell(g, (double)xa, (double)ya, (double)za, cx, cy, cz);
        dl = stg_reach<T>(ha, rs, g.cs);
    }
    const int dmax = stg_wave_max(dl);
    for (int s = 0; s < S.L.nsrc; ++s) {
        const T *srec = S.L.srec[s];
        const T *spm = S.L.spm[s];
        const int32_t *row = S.L.indices[s] + S.L.slice_off[s][slice] + lane;
        T nr = T(0.0);
        auto stage = [&](P *dst, int32_t p) {
            const P *sp = pm2(spm, srec, p);
            dst[0] = sp[0];
            dst[1] = sp[1];
            if (!UH) dst[2] = P{srec[(int64_t)p * SPH_REC + SPH_R_H], T(0)};
        };
        auto fetch = [&](P *pr, int32_t p) { stage(pr, p); };
        auto term = [&](const P *pr) {
            nr += rho_term<UH, T>(dim, xa, ya, za, ha, facA, pr[0], pr[1],
                                  UH ? ha : pr[NP - 1].x, invA);
        };
        stg_row<T, NP, CAP>(g, cx, cy, cz, dmax, S.sstart[s], row, e[s], (int32_t)t,
                            S.same[s] != 0, lds, stage, fetch, term);
        if (any) val = val + nr;
    }
    if (!real) {
        if (act) {  // non-LocalReal destinations (sph_func.pyx:226-227)
            rho[i] = 0;
            if (out_s) out_s[t * out_stride] = 0;
        }
        return;
    }
    if (!any) val = val + T(0.0);
    rho[i] = val;
    if (out_s) out_s[t * out_stride] = val;
}

// ---------------------------------------------------------------------------
// MomentumEquation + artificial viscosity (pressure_funcs.pyx:246-391; the
// arithmetic of k_momentum)
// ---------------------------------------------------------------------------
template <bool UH, typename T, int CAP>
__global__ void __launch_bounds__(STG_THREADS)
    k_momentum_staged(int dim, int64_t nd_total, int64_t nd_real,
                      const int32_t *__restrict__ dperm, StagedSrc<T> S,
                      const T *__restrict__ drec, const double *__restrict__ ox,
                      const double *__restrict__ oy, const double *__restrict__ oz, GridDev g,
                      double rs, T alpha, T beta, T eta, int accumulate,
                      double *__restrict__ ax, double *__restrict__ ay, double *__restrict__ az,
                      double *__restrict__ dt_fac)
{
    using P = typename Pair<T>::type;
    constexpr int NP = UH ? 5 : 6;  // (x, y) (z, m) (u, v) (w, P/rho^2) (rho, cs) [(h, p)]
    __shared__ P lds[CAP * NP];
    const int lane = threadIdx.x;
    const int64_t slice = sph_xcd_tile(blockIdx.x, gridDim.x);
    const int64_t t = slice * STG_THREADS + lane;
    const bool act = t < nd_total;
    const int32_t i = act ? dperm[t] : 0;
    const bool real = act && i < nd_real;
    int32_t e[SPH_MAX_SRC];
    int32_t any = 0;
#pragma unroll
    for (int s = 0; s < SPH_MAX_SRC; ++s) {
        e[s] = (real && s < S.L.nsrc) ? S.L.counts[s][t] : 0;
        any |= e[s];
    }
    T dtf = T(0), o0 = T(0), o1 = T(0), o2 = T(0);
    if (real) {
        dtf = T(dt_fac[i]);
        o0 = accumulate ? T(ax[i]) : T(0.0);
        o1 = accumulate ? T(ay[i]) : T(0.0);
        o2 = (accumulate && az) ? T(az[i]) : T(0.0);
    }
    MomDest<T> D;
    memset(&D, 0, sizeof(D));
    int cx = 0, cy = 0, cz = 0, dl = 0;
    if (any) {
        const auto *dr = rec2(drec, t);
        const auto d0 = dr[0], d1 = dr[1], d2 = dr[2], d3 = dr[3], d4 = dr[4], d5 = dr[5];
        D.xa = d0.x;
        D.ya = d0.y;
        D.za = d1.x;
        D.ua = d2.x;
        D.va = d2.y;
        D.wa = d3.x;
        D.rhoa = d4.x;
        D.ca = d4.y;
        D.ha = d5.x;
        D.Pa = d5.y;
        D.facA = sph_h_dim(D.ha, dim) * sph_cubic_fac1<T>(dim);
        D.inv_ha = sph_inv_for_div(D.ha);
        D.rhoa2 = T(1.0) / (D.rhoa * D.rhoa);
        if (ox) stg_cell(g, ox[i], oy[i], oz[i], cx, cy, cz);
        else stg_cell(g, (double)D.xa, (double)D.ya, (double)D.za, cx, cy, cz);
        dl = stg_reach<T>(D.ha, rs, g.cs);
    }
    const int dmax = stg_wave_max(dl);
    for (int s = 0; s < S.L.nsrc; ++s) {
        const T *srec = S.L.srec[s];
        const int32_t *row = S.L.indices[s] + S.L.slice_off[s][slice] + lane;
        T nr0 = T(0.0), nr1 = T(0.0), nr2 = T(0.0);
        auto stage = [&](P *dst, int32_t p) {
            const P *sp = rec2(srec, p);
#pragma unroll
            for (int f = 0; f < NP; ++f) dst[f] = sp[f];
        };
        auto fetch = [&](P *pr, int32_t p) { stage(pr, p); };
        auto term = [&](const P *pr) {
            mom_term<UH, T>(dim, D, alpha, beta, eta, pr[0], pr[1], pr[2], pr[3], pr[4],
                            UH ? D.ha : pr[NP - 1].x, dtf, nr0, nr1, nr2);
        };
        stg_row<T, NP, CAP>(g, cx, cy, cz, dmax, S.sstart[s], row, e[s], (int32_t)t,
                            S.same[s] != 0, lds, stage, fetch, term);
        if (any) {
            o0 = o0 + nr0;
            o1 = o1 + nr1;
            o2 = o2 + nr2;
        }
    }
    if (!real) {
        if (act) {
            ax[i] = 0;
            ay[i] = 0;
            if (az) az[i] = 0;
        }
        return;
    }
    if (!any) {  // empty rows: dt_fac keeps its value, a = base + 0
        o0 = o0 + T(0.0);
        o1 = o1 + T(0.0);
        o2 = o2 + T(0.0);
    }
    dt_fac[i] = dtf;
    ax[i] = o0;
    ay[i] = o1;
    if (az) az[i] = o2;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
#ifndef STG_DENS_CAP
#define STG_DENS_CAP 256
#endif
#ifndef STG_MOM_CAP
#define STG_MOM_CAP 192
#endif

template <typename T>
static int staged_sources(sph_ctx *ctx, const char *name, const sph_grid *grid,
                          const sph_sources *src, StagedSrc<T> *S, GridDev *g)
{
    SPH_REQUIRE(ctx, grid && grid->order == SPH_ORDER_LINEAR,
                "%s: the staged kernels need the linear cell order", name);
    int rc = src_lists<T>(ctx, name, src, &S->L);
    if (rc) return rc;
    for (int s = 0; s < src->nsrc; ++s) {
        SPH_REQUIRE(ctx, src->s_cell_start[s], "%s: null cell starts of source %d", name, s);
        S->sstart[s] = src->s_cell_start[s];
        S->same[s] = src->same[s];
    }
    rc = sph_upload_key_tables(ctx, grid);
    if (rc) return rc;
    g->cs = grid->cell_size;
    g->cmin0 = grid->cmin[0];
    g->cmin1 = grid->cmin[1];
    g->cmin2 = grid->cmin[2];
    g->d0 = grid->dims[0];
    g->d1 = grid->dims[1];
    g->d2 = grid->dims[2];
    g->tab = ctx->key_tab;
    g->key_base = 0;
    return SPH_OK;
}

template <typename T>
static int density_staged(sph_ctx *ctx, const sph_grid *grid, double radius_scale, int dim,
                          int64_t nd_total, int64_t nd_real, const int32_t *d_perm,
                          const sph_sources *src, const T *d_rec, const double *ox,
                          const double *oy, const double *oz, int uniform_h, int accumulate,
                          double *rho, T *out_s, int64_t out_s_stride)
{
    if (!ctx) return SPH_EINVAL;
    const char *name = "sph_summation_density_staged";
    SPH_REQUIRE(ctx, dim >= 1 && dim <= 3, "%s: bad dim %d", name, dim);
    SPH_REQUIRE(ctx, nd_real >= 0 && nd_total >= nd_real && nd_total < ((int64_t)1 << 31),
                "%s: bad n", name);
    if (nd_total == 0) return SPH_OK;
    SPH_REQUIRE(ctx, rho && d_perm && d_rec && (!out_s || out_s_stride >= 1),
                "%s: null argument", name);
    StagedSrc<T> S;
    memset(&S, 0, sizeof(S));
    GridDev g;
    int rc = staged_sources<T>(ctx, name, grid, src, &S, &g);
    if (rc) return rc;
    auto kern = uniform_h ? k_density_staged<true, T, STG_DENS_CAP>
                          : k_density_staged<false, T, STG_DENS_CAP>;
    hipLaunchKernelGGL(kern, dim3(sph_blocks(nd_total, STG_THREADS)), dim3(STG_THREADS), 0,
                       ctx->stream, dim, nd_total, nd_real, d_perm, S, d_rec, ox, oy, oz, g,
                       radius_scale, accumulate, rho, out_s, out_s_stride);
    SPH_LAUNCH_CHECK(ctx, "k_density_staged");
    return SPH_OK;
}

template <typename T>
static int momentum_staged(sph_ctx *ctx, const sph_grid *grid, double radius_scale, int dim,
                           int64_t nd_total, int64_t nd_real, const int32_t *d_perm,
                           const sph_sources *src, const T *d_rec, const double *ox,
                           const double *oy, const double *oz, int uniform_h, double alpha,
                           double beta, double eta, int accumulate, double *ax, double *ay,
                           double *az, double *dt_fac)
{
    if (!ctx) return SPH_EINVAL;
    const char *name = "sph_momentum_av_staged";
    SPH_REQUIRE(ctx, dim >= 1 && dim <= 3, "%s: bad dim %d", name, dim);
    SPH_REQUIRE(ctx, nd_real >= 0 && nd_total >= nd_real && nd_total < ((int64_t)1 << 31),
                "%s: bad n", name);
    if (nd_total == 0) return SPH_OK;
    SPH_REQUIRE(ctx, d_rec && ax && ay && dt_fac && d_perm, "%s: null argument", name);
    StagedSrc<T> S;
    memset(&S, 0, sizeof(S));
    GridDev g;
    int rc = staged_sources<T>(ctx, name, grid, src, &S, &g);
    if (rc) return rc;
    rc = sph_prho2_sources(ctx, src, sizeof(T) == sizeof(float));
    if (rc) return rc;
    auto kern = uniform_h ? k_momentum_staged<true, T, STG_MOM_CAP>
                          : k_momentum_staged<false, T, STG_MOM_CAP>;
    hipLaunchKernelGGL(kern, dim3(sph_blocks(nd_total, STG_THREADS)), dim3(STG_THREADS), 0,
                       ctx->stream, dim, nd_total, nd_real, d_perm, S, d_rec, ox, oy, oz, g,
                       radius_scale, T(alpha), T(beta), T(eta), accumulate, ax, ay, az, dt_fac);
    SPH_LAUNCH_CHECK(ctx, "k_momentum_staged");
    return SPH_OK;
}

extern "C" int sph_summation_density_staged(sph_ctx *ctx, const sph_grid *grid,
                                            double radius_scale, int dim, int64_t nd_total,
                                            int64_t nd_real, const int32_t *d_perm,
                                            const sph_sources *src, const double *d_rec,
                                            int uniform_h, int accumulate, double *rho,
                                            double *out_s, int64_t out_s_stride)
{
    return density_staged<double>(ctx, grid, radius_scale, dim, nd_total, nd_real, d_perm, src,
                                  d_rec, nullptr, nullptr, nullptr, uniform_h, accumulate, rho,
                                  out_s, out_s_stride);
}

extern "C" int sph_momentum_av_staged(sph_ctx *ctx, const sph_grid *grid, double radius_scale,
                                      int dim, int64_t nd_total, int64_t nd_real,
                                      const int32_t *d_perm, const sph_sources *src,
                                      const double *d_rec, int uniform_h, double alpha,
                                      double beta, double eta, int accumulate, double *ax,
                                      double *ay, double *az, double *dt_fac)
{
    return momentum_staged<double>(ctx, grid, radius_scale, dim, nd_total, nd_real, d_perm, src,
                                   d_rec, nullptr, nullptr, nullptr, uniform_h, alpha, beta, eta,
                                   accumulate, ax, ay, az, dt_fac);
}

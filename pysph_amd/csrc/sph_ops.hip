// SPH operators over neighbour lists and the integrator stage arithmetic.
//
// Every per-pair expression follows the reference Cython in order
// (sph/funcs/density_funcs.pyx:37-73, sph/funcs/eos_funcs.pyx:118-129,
// sph/funcs/pressure_funcs.pyx:246-391) and is compiled without FMA
// contraction, so with the lists in the reference's order the sums are the
// reference's sums.
#include <string.h>

#include "sph_internal.h"
#include "sph_terms.h"

#define OP_THREADS 256
// neighbour records whose loads are in flight before their terms are added
// (measured on the 100^3 lattice with 75 neighbours per particle: density
// 0.34 ms at 4, 0.31 ms at 8; momentum 0.66 ms at 2, 0.60 ms at 4; on the
// 1M dam break, density with the Tait equation fused: 0.34 ms at 8 (112
// registers, 4 waves per SIMD), 0.35 ms at 12 (3 waves), 0.41 ms at 16)
#ifndef DENS_NT
#define DENS_NT 8
#endif
#ifndef MOM_NT
#define MOM_NT 4
#endif
// waves per SIMD the momentum kernel is compiled for: 3 caps it at 168
// registers (172 unbounded, which allows 2 waves), with two 8-byte spills
// outside the neighbour loop; 0.657 -> 0.649 ms on the 1M dam break
#ifndef MOM_WAVES
#define MOM_WAVES 3
#endif

// ---------------------------------------------------------------------------
// Tait equation of state of one particle (eos_funcs.pyx:118-129)
// ---------------------------------------------------------------------------
template <typename T> struct TaitArgs {
    T ro, co, gamma, B;  // B = co^2 ro / gamma (eos_funcs.pyx:90)
    int accumulate;
    double *p, *cs;      // original order
    T *p_s, *cs_s, *prho2_s;  // sorted record fields (or NULL), stride st
    int64_t st;
};

// p, cs of sorted particle t (original index i) from its density rhoa (the
// record's value); a non-LocalReal particle gets 0 (sph_func.pyx:226-227).
// P/rho^2 of the record, k_prho2's two operations, when prho2_s is given.
template <typename T>
__device__ __forceinline__ void tait_point(const TaitArgs<T> &E, int64_t t, int32_t i,
                                           bool real, T rhoa)
{
    T r0 = T(0.0), r1 = T(0.0);
    if (real) {
        const T ratio = rhoa / E.ro;
        const T gamma2 = T(0.5) * (E.gamma - T(1.0));
        const T tmp = sph_pow(ratio, E.gamma);
        r0 = (tmp - T(1.0)) * E.B;
        r1 = sph_pow(ratio, gamma2) * E.co;
        if (E.accumulate) {
            r0 = (E.p_s ? E.p_s[t * E.st] : T(E.p[i])) + r0;
            r1 = (E.cs_s ? E.cs_s[t * E.st] : T(E.cs[i])) + r1;
        } else {
            r0 = T(0.0) + r0;
            r1 = T(0.0) + r1;
        }
    }
    E.p[i] = r0;
    E.cs[i] = r1;
    if (E.p_s) E.p_s[t * E.st] = r0;
    if (E.cs_s) E.cs_s[t * E.st] = r1;
    if (E.prho2_s) {
        const T rhob2 = T(1.0) / (rhoa * rhoa);
        E.prho2_s[t * E.st] = r0 * rhob2;
    }
}

// ---------------------------------------------------------------------------
// SPHRho (rho_term in sph_terms.h)
// ---------------------------------------------------------------------------
// one SELL-64 row of SPHRho, summed from 0 in list order
template <bool UH, typename T>
__device__ __forceinline__ T rho_row(int dim, T xa, T ya, T za, T ha, T facA,
                                     const int32_t *__restrict__ row, int32_t e,
                                     const T *__restrict__ srec, const T *__restrict__ spm)
{
    T nr = T(0.0);
    int32_t k = 0;
    const T invA = sph_inv_for_div(ha);
    // DENS_NT neighbours' loads in flight, then their terms added in list order;
    // each list load is one coalesced line for the wavefront, read once, so it
    // bypasses the caches
    for (; k + DENS_NT <= e; k += DENS_NT) {
        int32_t j[DENS_NT];
        typename Pair<T>::type a[DENS_NT], b[DENS_NT];
        T hb[DENS_NT];
#pragma unroll
        for (int u = 0; u < DENS_NT; ++u)
            j[u] = __builtin_nontemporal_load(row + ((int64_t)(k + u) << 6));
#pragma unroll
        for (int u = 0; u < DENS_NT; ++u) {
            const auto *sp = pm2(spm, srec, j[u]);
            a[u] = sp[0];
            b[u] = sp[1];
            hb[u] = UH ? ha : srec[(int64_t)j[u] * SPH_REC + SPH_R_H];
        }
#pragma unroll
        for (int u = 0; u < DENS_NT; ++u)
            nr += rho_term<UH, T>(dim, xa, ya, za, ha, facA, a[u], b[u], hb[u], invA);
    }
    for (; k < e; ++k) {
        const int32_t j = __builtin_nontemporal_load(row + ((int64_t)k << 6));
        const auto *sp = pm2(spm, srec, j);
        const T hb = UH ? ha : srec[(int64_t)j * SPH_REC + SPH_R_H];
        nr += rho_term<UH, T>(dim, xa, ya, za, ha, facA, sp[0], sp[1], hb, invA);
    }
    return nr;
}

// EOS: TaitEquation of the same particle follows, from the density just
// written to its record (the k_tait arithmetic; the caller passes out_s)
template <bool UH, bool EOS, typename T>
__global__ void __launch_bounds__(OP_THREADS)
    k_summation_density(int dim, int64_t nd_total, int64_t nd_real,
                        const int32_t *__restrict__ dperm, SrcLists<T> L,
                        const T *__restrict__ drec, int accumulate,
                        double *__restrict__ rho, T *__restrict__ out_s, int64_t out_stride,
                        TaitArgs<T> E)
{
    const int64_t t =
        (int64_t)sph_xcd_tile(blockIdx.x, gridDim.x) * OP_THREADS + threadIdx.x;
    if (t >= nd_total) return;
    const int32_t i = dperm[t];
    if (i >= nd_real) {  // non-LocalReal destinations (sph_func.pyx:226-227)
        rho[i] = 0;
        if (out_s) out_s[t * out_stride] = 0;
        // the record's density as k_tait reads it (a runtime value: a
        // constant 0 would let the compiler fold P/rho^2's 0 * inf)
        if (EOS) tait_point(E, t, i, false, *(volatile T *)&out_s[t * out_stride]);
        return;
    }
    int32_t e[SPH_MAX_SRC];
    int32_t any = 0;
#pragma unroll
    for (int s = 0; s < SPH_MAX_SRC; ++s) {
        e[s] = s < L.nsrc ? L.counts[s][t] : 0;
        any |= e[s];
    }
    T val = accumulate ? (out_s ? out_s[t * out_stride] : T(rho[i])) : T(0.0);
    if (!any) {  // empty rows (a destination far from a second array): val + 0
        val = val + T(0.0);
    } else {
        const auto *dr = rec2(drec, t);
        const auto dxy = dr[0], dzm = dr[1];
        const T xa = dxy.x, ya = dxy.y, za = dzm.x, ha = drec[t * SPH_REC + SPH_R_H];
        const T facA = sph_h_dim(ha, dim) * sph_cubic_fac1<T>(dim);
        for (int s = 0; s < L.nsrc; ++s) {
            const int32_t *row = L.indices[s] + L.slice_off[s][t >> 6] + (t & 63);
            val = val + rho_row<UH, T>(dim, xa, ya, za, ha, facA, row, e[s], L.srec[s],
                                       L.spm[s]);
        }
    }
    rho[i] = val;
    if (out_s) out_s[t * out_stride] = val;
    if (EOS) tait_point(E, t, i, true, val);
}

template <typename T>
static int summation_density(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                             const int32_t *d_perm, const sph_sources *src, const T *d_rec,
                             int uniform_h, int accumulate, double *rho, T *out_s,
                             int64_t out_s_stride, const TaitArgs<T> *eos = nullptr)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, dim >= 1 && dim <= 3, "sph_summation_density: bad dim %d", dim);
    SPH_REQUIRE(ctx, nd_real >= 0 && nd_total >= nd_real, "sph_summation_density: bad n");
    if (nd_total == 0) return SPH_OK;
    SPH_REQUIRE(ctx, rho && d_perm && d_rec && (!out_s || out_s_stride >= 1),
                "sph_summation_density: null argument");
    SrcLists<T> L;
    int rc = src_lists<T>(ctx, "sph_summation_density", src, &L);
    if (rc) return rc;
    TaitArgs<T> E;
    memset(&E, 0, sizeof(E));
    if (eos) E = *eos;
    auto kern = eos ? (uniform_h ? k_summation_density<true, true, T>
                                 : k_summation_density<false, true, T>)
                    : (uniform_h ? k_summation_density<true, false, T>
                                 : k_summation_density<false, false, T>);
    hipLaunchKernelGGL(kern, dim3(sph_blocks(nd_total, OP_THREADS)),
                       dim3(OP_THREADS), 0, ctx->stream, dim, nd_total, nd_real, d_perm, L,
                       d_rec, accumulate, rho, out_s, out_s_stride, E);
    SPH_LAUNCH_CHECK(ctx, "k_summation_density");
    return SPH_OK;
}

static sph_sources one_source(const int32_t *counts, const int64_t *slice_off,
                              const int32_t *indices, const void *s_rec, int64_t ns)
{
    sph_sources src;
    memset(&src, 0, sizeof(src));
    src.nsrc = 1;
    src.counts[0] = counts;
    src.slice_off[0] = slice_off;
    src.indices[0] = indices;
    src.s_rec[0] = s_rec;
    src.ns[0] = ns;
    return src;
}

extern "C" int sph_summation_density(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                                     const int32_t *d_perm, const int32_t *counts,
                                     const int64_t *slice_off, const int32_t *indices,
                                     const double *d_rec, const double *s_rec, int uniform_h,
                                     int accumulate, double *rho, double *out_s,
                                     int64_t out_s_stride)
{
    const sph_sources src = one_source(counts, slice_off, indices, s_rec, 0);
    return summation_density<double>(ctx, dim, nd_total, nd_real, d_perm, &src, d_rec,
                                      uniform_h, accumulate, rho, out_s, out_s_stride);
}

extern "C" int sph_summation_density_f32(sph_ctx *ctx, int dim, int64_t nd_total,
                                         int64_t nd_real, const int32_t *d_perm,
                                         const int32_t *counts, const int64_t *slice_off,
                                         const int32_t *indices, const float *d_rec,
                                         const float *s_rec, int uniform_h, int accumulate,
                                         double *rho, float *out_s, int64_t out_s_stride)
{
    const sph_sources src = one_source(counts, slice_off, indices, s_rec, 0);
    return summation_density<float>(ctx, dim, nd_total, nd_real, d_perm, &src, d_rec,
                                     uniform_h, accumulate, rho, out_s, out_s_stride);
}

extern "C" int sph_summation_density_multi(sph_ctx *ctx, int dim, int64_t nd_total,
                                           int64_t nd_real, const int32_t *d_perm,
                                           const sph_sources *src, const double *d_rec,
                                           int uniform_h, int accumulate, double *rho,
                                           double *out_s, int64_t out_s_stride)
{
    return summation_density<double>(ctx, dim, nd_total, nd_real, d_perm, src, d_rec,
                                      uniform_h, accumulate, rho, out_s, out_s_stride);
}

extern "C" int sph_summation_density_multi_f32(sph_ctx *ctx, int dim, int64_t nd_total,
                                               int64_t nd_real, const int32_t *d_perm,
                                               const sph_sources *src, const float *d_rec,
                                               int uniform_h, int accumulate, double *rho,
                                               float *out_s, int64_t out_s_stride)
{
    return summation_density<float>(ctx, dim, nd_total, nd_real, d_perm, src, d_rec,
                                     uniform_h, accumulate, rho, out_s, out_s_stride);
}

// ---------------------------------------------------------------------------
// Tait equation of state
// ---------------------------------------------------------------------------
template <typename T>
static TaitArgs<T> tait_args(double ro, double co, double gamma, int accumulate, double *p,
                             double *cs, T *p_s, T *cs_s, T *prho2_s, int64_t st)
{
    TaitArgs<T> E;
    E.ro = T(ro);
    E.co = T(co);
    E.gamma = T(gamma);
    E.B = T(co) * T(co) * T(ro) / T(gamma);  // eos_funcs.pyx:90
    E.accumulate = accumulate;
    E.p = p;
    E.cs = cs;
    E.p_s = p_s;
    E.cs_s = cs_s;
    E.prho2_s = prho2_s;
    E.st = st;
    return E;
}

template <typename T>
__global__ void __launch_bounds__(OP_THREADS)
    k_tait(int64_t n_total, int64_t n_real, const int32_t *__restrict__ perm,
           const T *__restrict__ rho_s, int64_t st, TaitArgs<T> E)
{
    int64_t t = (int64_t)blockIdx.x * OP_THREADS + threadIdx.x;
    if (t >= n_total) return;
    const int32_t i = perm[t];
    tait_point(E, t, i, i < n_real, rho_s[t * st]);
}

template <typename T>
static int tait(sph_ctx *ctx, int64_t n_total, int64_t n_real, const int32_t *perm,
                const T *rho_s, int64_t s_stride, double ro, double co, double gamma,
                int accumulate, double *p, double *cs, T *p_s, T *cs_s, T *prho2_s)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, n_real >= 0 && n_total >= n_real, "sph_tait: bad n");
    if (n_total == 0) return SPH_OK;
    SPH_REQUIRE(ctx, perm && rho_s && p && cs && s_stride >= 1, "sph_tait: null argument");
    const TaitArgs<T> E = tait_args<T>(ro, co, gamma, accumulate, p, cs, p_s, cs_s, prho2_s,
                                       s_stride);
    hipLaunchKernelGGL(k_tait<T>, dim3(sph_blocks(n_total, OP_THREADS)), dim3(OP_THREADS), 0,
                       ctx->stream, n_total, n_real, perm, rho_s, s_stride, E);
    SPH_LAUNCH_CHECK(ctx, "k_tait");
    return SPH_OK;
}

extern "C" int sph_tait(sph_ctx *ctx, int64_t n_total, int64_t n_real, const int32_t *perm,
                        const double *rho_s, int64_t s_stride, double ro, double co,
                        double gamma, int accumulate, double *p, double *cs, double *p_s,
                        double *cs_s, double *prho2_s)
{
    return tait<double>(ctx, n_total, n_real, perm, rho_s, s_stride, ro, co, gamma, accumulate,
                        p, cs, p_s, cs_s, prho2_s);
}

extern "C" int sph_tait_f32(sph_ctx *ctx, int64_t n_total, int64_t n_real, const int32_t *perm,
                            const float *rho_s, int64_t s_stride, double ro, double co,
                            double gamma, int accumulate, double *p, double *cs, float *p_s,
                            float *cs_s, float *prho2_s)
{
    return tait<float>(ctx, n_total, n_real, perm, rho_s, s_stride, ro, co, gamma, accumulate,
                       p, cs, p_s, cs_s, prho2_s);
}

// SPHRho followed by TaitEquation on the same destination array in one launch
// (the density sums of sph_summation_density_multi, then k_tait's arithmetic
// on the density just written to the record).  The Tait record fields share
// the density record's stride; out_s (the record's rho) is required.
template <typename T>
static int summation_density_tait(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                                  const int32_t *d_perm, const sph_sources *src, const T *d_rec,
                                  int uniform_h, int accumulate, double *rho, T *out_s,
                                  int64_t out_s_stride, double ro, double co, double gamma,
                                  int eos_accumulate, double *p, double *cs, T *p_s, T *cs_s,
                                  T *prho2_s)
{
    if (!ctx) return SPH_EINVAL;
    if (nd_total == 0) return SPH_OK;
    SPH_REQUIRE(ctx, out_s && p && cs, "sph_summation_density_tait: null argument");
    const TaitArgs<T> E = tait_args<T>(ro, co, gamma, eos_accumulate, p, cs, p_s, cs_s,
                                       prho2_s, out_s_stride);
    return summation_density<T>(ctx, dim, nd_total, nd_real, d_perm, src, d_rec, uniform_h,
                                accumulate, rho, out_s, out_s_stride, &E);
}

extern "C" int sph_summation_density_tait(sph_ctx *ctx, int dim, int64_t nd_total,
                                          int64_t nd_real, const int32_t *d_perm,
                                          const sph_sources *src, const double *d_rec,
                                          int uniform_h, int accumulate, double *rho,
                                          double *out_s, int64_t out_s_stride, double ro,
                                          double co, double gamma, int eos_accumulate,
                                          double *p, double *cs, double *p_s, double *cs_s,
                                          double *prho2_s)
{
    return summation_density_tait<double>(ctx, dim, nd_total, nd_real, d_perm, src, d_rec,
                                          uniform_h, accumulate, rho, out_s, out_s_stride, ro,
                                          co, gamma, eos_accumulate, p, cs, p_s, cs_s, prho2_s);
}

extern "C" int sph_summation_density_tait_f32(sph_ctx *ctx, int dim, int64_t nd_total,
                                              int64_t nd_real, const int32_t *d_perm,
                                              const sph_sources *src, const float *d_rec,
                                              int uniform_h, int accumulate, double *rho,
                                              float *out_s, int64_t out_s_stride, double ro,
                                              double co, double gamma, int eos_accumulate,
                                              double *p, double *cs, float *p_s, float *cs_s,
                                              float *prho2_s)
{
    return summation_density_tait<float>(ctx, dim, nd_total, nd_real, d_perm, src, d_rec,
                                         uniform_h, accumulate, rho, out_s, out_s_stride, ro,
                                         co, gamma, eos_accumulate, p, cs, p_s, cs_s, prho2_s);
}

// Tait over sorted records with every particle real and no accumulation
// (the batched path): p, cs to the original-order arrays and to the
// record, with P/rho^2 (the k_prho2 value) written in the same pass
template <typename T>
__global__ void __launch_bounds__(OP_THREADS)
    k_tait_rec(int64_t n, const int32_t *__restrict__ perm, T *__restrict__ rec, T ro, T co,
               T gamma, T B, double *__restrict__ p, double *__restrict__ cs)
{
    const int64_t t = (int64_t)blockIdx.x * OP_THREADS + threadIdx.x;
    if (t >= n) return;
    T *r = rec + t * SPH_REC;
    const T rhoa = r[SPH_R_RHO];
    const T ratio = rhoa / ro;
    const T gamma2 = T(0.5) * (gamma - T(1.0));
    const T tmp = sph_pow(ratio, gamma);
    const T r0 = T(0.0) + (tmp - T(1.0)) * B;
    const T r1 = T(0.0) + sph_pow(ratio, gamma2) * co;
    const int32_t i = perm[t];
    p[i] = r0;
    cs[i] = r1;
    r[SPH_R_P] = r0;
    r[SPH_R_CS] = r1;
    const T rhob2 = T(1.0) / (rhoa * rhoa);
    r[SPH_R_PRHO2] = r0 * rhob2;
}

template <typename T>
static int tait_records(sph_ctx *ctx, int64_t n, const int32_t *perm, T *rec, double ro,
                        double co, double gamma, double *p, double *cs)
{
    if (n == 0) return SPH_OK;
    const T B = T(co) * T(co) * T(ro) / T(gamma);  // eos_funcs.pyx:90
    hipLaunchKernelGGL(k_tait_rec<T>, dim3(sph_blocks(n, OP_THREADS)), dim3(OP_THREADS), 0,
                       ctx->stream, n, perm, rec, T(ro), T(co), T(gamma), B, p, cs);
    SPH_LAUNCH_CHECK(ctx, "k_tait_rec");
    return SPH_OK;
}

int sph_tait_records(sph_ctx *ctx, int64_t n, const int32_t *perm, double *rec, double ro,
                     double co, double gamma, double *p, double *cs)
{
    return tait_records<double>(ctx, n, perm, rec, ro, co, gamma, p, cs);
}

int sph_tait_records_f32(sph_ctx *ctx, int64_t n, const int32_t *perm, float *rec, double ro,
                         double co, double gamma, double *p, double *cs)
{
    return tait_records<float>(ctx, n, perm, rec, ro, co, gamma, p, cs);
}

// ---------------------------------------------------------------------------
// MomentumEquation with Monaghan artificial viscosity (mom_term in sph_terms.h)
// ---------------------------------------------------------------------------
// one SELL-64 row of MomentumEquation, summed from 0 in list order; dtf is
// max-reduced over the row.  NT neighbours' records in flight.
template <bool UH, typename T, int NT>
__device__ __forceinline__ void mom_row(int dim, const MomDest<T> &D, T alpha, T beta, T eta,
                                        const int32_t *__restrict__ row, int32_t e,
                                        const T *__restrict__ srec, T &dtf, T &nr0, T &nr1,
                                        T &nr2)
{
    int32_t k = 0;
    // NT neighbours' records in flight, then their terms added in list order
    for (; k + NT <= e; k += NT) {
        int32_t j[NT];
        typename Pair<T>::type s0[NT], s1[NT], s2[NT], s3[NT], s4[NT];
        T hb[NT];
#pragma unroll
        for (int u = 0; u < NT; ++u)
            j[u] = __builtin_nontemporal_load(row + ((int64_t)(k + u) << 6));
#pragma unroll
        for (int u = 0; u < NT; ++u) {
            const auto *sr = rec2(srec, j[u]);
            s0[u] = sr[0];
            s1[u] = sr[1];
            s2[u] = sr[2];
            s3[u] = sr[3];
            s4[u] = sr[4];
            hb[u] = UH ? D.ha : srec[(int64_t)j[u] * SPH_REC + SPH_R_H];
        }
#pragma unroll
        for (int u = 0; u < NT; ++u)
            mom_term<UH, T>(dim, D, alpha, beta, eta, s0[u], s1[u], s2[u], s3[u], s4[u], hb[u],
                            dtf, nr0, nr1, nr2);
    }
    for (; k < e; ++k) {
        const int32_t j = __builtin_nontemporal_load(row + ((int64_t)k << 6));
        const auto *sr = rec2(srec, j);
        const T hb = UH ? D.ha : srec[(int64_t)j * SPH_REC + SPH_R_H];
        mom_term<UH, T>(dim, D, alpha, beta, eta, sr[0], sr[1], sr[2], sr[3], sr[4], hb, dtf,
                        nr0, nr1, nr2);
    }
}

template <bool UH, typename T>
__global__ void __launch_bounds__(OP_THREADS) __attribute__((amdgpu_waves_per_eu(MOM_WAVES)))
    k_momentum(int dim, int64_t nd_total, int64_t nd_real, const int32_t *__restrict__ dperm,
               SrcLists<T> L, const T *__restrict__ drec, T alpha, T beta, T eta,
               int accumulate, double *__restrict__ ax, double *__restrict__ ay,
               double *__restrict__ az, double *__restrict__ dt_fac)
{
    const int64_t t =
        (int64_t)sph_xcd_tile(blockIdx.x, gridDim.x) * OP_THREADS + threadIdx.x;
    if (t >= nd_total) return;
    const int32_t i = dperm[t];
    if (i >= nd_real) {
        ax[i] = 0;
        ay[i] = 0;
        if (az) az[i] = 0;
        return;
    }
    int32_t e[SPH_MAX_SRC];
    int32_t any = 0;
#pragma unroll
    for (int s = 0; s < SPH_MAX_SRC; ++s) {
        e[s] = s < L.nsrc ? L.counts[s][t] : 0;
        any |= e[s];
    }
    T dtf = T(dt_fac[i]);
    T o0 = accumulate ? T(ax[i]) : T(0.0);
    T o1 = accumulate ? T(ay[i]) : T(0.0);
    T o2 = (accumulate && az) ? T(az[i]) : T(0.0);
    if (!any) {  // empty rows: dt_fac keeps its value, a = base + 0
        o0 = o0 + T(0.0);
        o1 = o1 + T(0.0);
        o2 = o2 + T(0.0);
    } else {
        const auto *dr = rec2(drec, t);
        const auto d0 = dr[0], d1 = dr[1], d2 = dr[2], d3 = dr[3], d4 = dr[4], d5 = dr[5];
        MomDest<T> D;
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
        for (int s = 0; s < L.nsrc; ++s) {
            T nr0 = T(0.0), nr1 = T(0.0), nr2 = T(0.0);
            const int32_t *row = L.indices[s] + L.slice_off[s][t >> 6] + (t & 63);
            // with h read per pair (!UH) four records in flight would spill
            // registers at 3 waves per SIMD: two
            mom_row<UH, T, (UH || sizeof(T) == sizeof(float)) ? MOM_NT : 2>(
                dim, D, alpha, beta, eta, row, e[s], L.srec[s], dtf, nr0, nr1, nr2);
            o0 = o0 + nr0;
            o1 = o1 + nr1;
            o2 = o2 + nr2;
        }
    }
    dt_fac[i] = dtf;
    ax[i] = o0;
    ay[i] = o1;
    if (az) az[i] = o2;
}

// P/rho^2 of every source record, the per-pair term of pressure_funcs.pyx:302-305
template <typename T> __global__ void __launch_bounds__(OP_THREADS) k_prho2(int64_t n, T *__restrict__ rec)
{
    const int64_t j = (int64_t)blockIdx.x * OP_THREADS + threadIdx.x;
    if (j >= n) return;
    T *r = rec + j * SPH_REC;
    const T rhob = r[SPH_R_RHO];
    const T rhob2 = T(1.0) / (rhob * rhob);
    r[SPH_R_PRHO2] = r[SPH_R_P] * rhob2;
}

int sph_prho2_sources(sph_ctx *ctx, const sph_sources *src, int f32)
{
    for (int s = 0; s < src->nsrc; ++s) {
        if (src->ns[s] == 0) continue;
        const dim3 grid(sph_blocks(src->ns[s], OP_THREADS));
        if (f32)
            hipLaunchKernelGGL(k_prho2<float>, grid, dim3(OP_THREADS), 0, ctx->stream,
                               src->ns[s], (float *)src->s_rec[s]);
        else
            hipLaunchKernelGGL(k_prho2<double>, grid, dim3(OP_THREADS), 0, ctx->stream,
                               src->ns[s], (double *)src->s_rec[s]);
        SPH_LAUNCH_CHECK(ctx, "k_prho2");
    }
    return SPH_OK;
}

template <typename T>
static int momentum_av(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                       const int32_t *d_perm, const sph_sources *src, const T *d_rec,
                       int uniform_h, double alpha, double beta, double eta, int accumulate,
                       double *ax, double *ay, double *az, double *dt_fac)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, dim >= 1 && dim <= 3, "sph_momentum_av: bad dim %d", dim);
    SPH_REQUIRE(ctx, nd_real >= 0 && nd_total >= nd_real, "sph_momentum_av: bad n");
    if (nd_total == 0) return SPH_OK;
    SPH_REQUIRE(ctx, d_rec && ax && ay && dt_fac && d_perm, "sph_momentum_av: null argument");
    SrcLists<T> L;
    int rc = src_lists<T>(ctx, "sph_momentum_av", src, &L);
    if (rc) return rc;
    rc = sph_prho2_sources(ctx, src, sizeof(T) == sizeof(float));
    if (rc) return rc;
    auto kern = uniform_h ? k_momentum<true, T> : k_momentum<false, T>;
    hipLaunchKernelGGL(kern, dim3(sph_blocks(nd_total, OP_THREADS)), dim3(OP_THREADS), 0,
                       ctx->stream, dim, nd_total, nd_real, d_perm, L, d_rec, T(alpha),
                       T(beta), T(eta), accumulate, ax, ay, az, dt_fac);
    SPH_LAUNCH_CHECK(ctx, "k_momentum");
    return SPH_OK;
}

extern "C" int sph_momentum_av(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                               const int32_t *d_perm, const int32_t *counts,
                               const int64_t *slice_off, const int32_t *indices,
                               const double *d_rec, double *s_rec,
                               int64_t ns, int uniform_h, double alpha, double beta, double eta,
                               int accumulate,
                               double *ax, double *ay, double *az, double *dt_fac)
{
    const sph_sources src = one_source(counts, slice_off, indices, s_rec, ns);
    return momentum_av<double>(ctx, dim, nd_total, nd_real, d_perm, &src, d_rec, uniform_h,
                               alpha, beta, eta, accumulate, ax, ay, az, dt_fac);
}

extern "C" int sph_momentum_av_f32(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                                   const int32_t *d_perm, const int32_t *counts,
                                   const int64_t *slice_off, const int32_t *indices,
                                   const float *d_rec, float *s_rec, int64_t ns, int uniform_h,
                                   double alpha, double beta, double eta, int accumulate,
                                   double *ax, double *ay, double *az, double *dt_fac)
{
    const sph_sources src = one_source(counts, slice_off, indices, s_rec, ns);
    return momentum_av<float>(ctx, dim, nd_total, nd_real, d_perm, &src, d_rec, uniform_h,
                              alpha, beta, eta, accumulate, ax, ay, az, dt_fac);
}

extern "C" int sph_momentum_av_multi(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                                     const int32_t *d_perm, const sph_sources *src,
                                     const double *d_rec, int uniform_h, double alpha,
                                     double beta, double eta, int accumulate, double *ax,
                                     double *ay, double *az, double *dt_fac)
{
    return momentum_av<double>(ctx, dim, nd_total, nd_real, d_perm, src, d_rec, uniform_h,
                               alpha, beta, eta, accumulate, ax, ay, az, dt_fac);
}

extern "C" int sph_momentum_av_multi_f32(sph_ctx *ctx, int dim, int64_t nd_total,
                                         int64_t nd_real, const int32_t *d_perm,
                                         const sph_sources *src, const float *d_rec,
                                         int uniform_h, double alpha, double beta, double eta,
                                         int accumulate, double *ax, double *ay, double *az,
                                         double *dt_fac)
{
    return momentum_av<float>(ctx, dim, nd_total, nd_real, d_perm, src, d_rec, uniform_h,
                              alpha, beta, eta, accumulate, ax, ay, az, dt_fac);
}

// ---------------------------------------------------------------------------
// the cubic spline itself
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(OP_THREADS)
    k_cubic_spline(int dim, int64_t n, const double *__restrict__ ax,
                   const double *__restrict__ ay, const double *__restrict__ az,
                   const double *__restrict__ bx, const double *__restrict__ by,
                   const double *__restrict__ bz, const double *__restrict__ h,
                   double *__restrict__ w, double *__restrict__ gx, double *__restrict__ gy,
                   double *__restrict__ gz)
{
    int64_t i = (int64_t)blockIdx.x * OP_THREADS + threadIdx.x;
    if (i >= n) return;
    const double rx = ax[i] - bx[i], ry = ay[i] - by[i], rz = az[i] - bz[i];
    if (w) w[i] = sph_cubic_w(dim, rx, ry, rz, h[i]);
    if (gx || gy || gz) {
        const double g = sph_cubic_gfac(dim, rx, ry, rz, h[i]);
        if (gx) gx[i] = rx * g;
        if (gy) gy[i] = ry * g;
        if (gz) gz[i] = rz * g;
    }
}

extern "C" int sph_cubic_spline(sph_ctx *ctx, int dim, int64_t n, const double *ax,
                                const double *ay, const double *az, const double *bx,
                                const double *by, const double *bz, const double *h, double *w,
                                double *gx, double *gy, double *gz)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, dim >= 1 && dim <= 3, "sph_cubic_spline: bad dim %d", dim);
    if (n <= 0) return SPH_OK;
    SPH_REQUIRE(ctx, ax && ay && az && bx && by && bz && h, "sph_cubic_spline: null input");
    hipLaunchKernelGGL(k_cubic_spline, dim3(sph_blocks(n, OP_THREADS)), dim3(OP_THREADS), 0,
                       ctx->stream, dim, n, ax, ay, az, bx, by, bz, h, w, gx, gy, gz);
    SPH_LAUNCH_CHECK(ctx, "k_cubic_spline");
    return SPH_OK;
}

// ---------------------------------------------------------------------------
// gravity, position stepping
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(OP_THREADS)
    k_gravity(int64_t n_real, int64_t n_total, double gx, double gy, double gz, int accumulate,
              double *__restrict__ ax, double *__restrict__ ay, double *__restrict__ az)
{
    int64_t i = (int64_t)blockIdx.x * OP_THREADS + threadIdx.x;
    if (i >= n_total) return;
    if (i >= n_real) {
        ax[i] = 0;
        ay[i] = 0;
        if (az) az[i] = 0;
        return;
    }
    ax[i] = (accumulate ? ax[i] : 0.0) + gx;
    ay[i] = (accumulate ? ay[i] : 0.0) + gy;
    if (az) az[i] = (accumulate ? az[i] : 0.0) + gz;
}

extern "C" int sph_gravity(sph_ctx *ctx, int64_t n_real, int64_t n_total, double gx, double gy,
                           double gz, int accumulate, double *ax, double *ay, double *az)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, n_real >= 0 && n_total >= n_real, "sph_gravity: bad n");
    if (n_total == 0) return SPH_OK;
    SPH_REQUIRE(ctx, ax && ay, "sph_gravity: null argument");
    hipLaunchKernelGGL(k_gravity, dim3(sph_blocks(n_total, OP_THREADS)), dim3(OP_THREADS), 0,
                       ctx->stream, n_real, n_total, gx, gy, gz, accumulate, ax, ay, az);
    SPH_LAUNCH_CHECK(ctx, "k_gravity");
    return SPH_OK;
}

__global__ void __launch_bounds__(OP_THREADS)
    k_position_stepping(int64_t n_real, int64_t n_total, const double *__restrict__ u,
                        const double *__restrict__ v, const double *__restrict__ w,
                        double *__restrict__ ax, double *__restrict__ ay, double *__restrict__ az)
{
    int64_t i = (int64_t)blockIdx.x * OP_THREADS + threadIdx.x;
    if (i >= n_total) return;
    const bool real = i < n_real;
    if (ax) ax[i] = real ? u[i] : 0.0;
    if (ay) ay[i] = real ? v[i] : 0.0;
    if (az) az[i] = real ? w[i] : 0.0;
}

extern "C" int sph_position_stepping(sph_ctx *ctx, int64_t n_real, int64_t n_total,
                                     const double *u, const double *v, const double *w,
                                     double *ax, double *ay, double *az)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, n_real >= 0 && n_total >= n_real, "sph_position_stepping: bad n");
    if (n_total == 0) return SPH_OK;
    SPH_REQUIRE(ctx, (!ax || u) && (!ay || v) && (!az || w), "sph_position_stepping: null");
    hipLaunchKernelGGL(k_position_stepping, dim3(sph_blocks(n_total, OP_THREADS)),
                       dim3(OP_THREADS), 0, ctx->stream, n_real, n_total, u, v, w, ax, ay, az);
    SPH_LAUNCH_CHECK(ctx, "k_position_stepping");
    return SPH_OK;
}

// ---------------------------------------------------------------------------
// integrator stages
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(OP_THREADS)
    k_stage_step(int64_t n, sph_step_props pr, double dt)
{
    int64_t i = (int64_t)blockIdx.x * OP_THREADS + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < pr.nprops; ++k) {
        const double t = pr.a[k][i] * dt;  // numpy: step_arr * dt, then + initial
        pr.x[k][i] = pr.x0[k][i] + t;
    }
}

__global__ void __launch_bounds__(OP_THREADS) k_pc_final(int64_t n, sph_step_props pr)
{
    int64_t i = (int64_t)blockIdx.x * OP_THREADS + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < pr.nprops; ++k) {
        const double t = 2 * pr.x[k][i];
        pr.x[k][i] = t - pr.x0[k][i];
    }
}

extern "C" int sph_stage_step(sph_ctx *ctx, int64_t n, const sph_step_props *props, double dt)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, props && props->nprops >= 0 && props->nprops <= SPH_MAX_PROPS,
                "sph_stage_step: bad props");
    if (n <= 0 || props->nprops == 0) return SPH_OK;
    hipLaunchKernelGGL(k_stage_step, dim3(sph_blocks(n, OP_THREADS)), dim3(OP_THREADS), 0,
                       ctx->stream, n, *props, dt);
    SPH_LAUNCH_CHECK(ctx, "k_stage_step");
    return SPH_OK;
}

extern "C" int sph_pc_final(sph_ctx *ctx, int64_t n, const sph_step_props *props)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, props && props->nprops >= 0 && props->nprops <= SPH_MAX_PROPS,
                "sph_pc_final: bad props");
    if (n <= 0 || props->nprops == 0) return SPH_OK;
    hipLaunchKernelGGL(k_pc_final, dim3(sph_blocks(n, OP_THREADS)), dim3(OP_THREADS), 0,
                       ctx->stream, n, *props);
    SPH_LAUNCH_CHECK(ctx, "k_pc_final");
    return SPH_OK;
}

__global__ void __launch_bounds__(OP_THREADS)
    k_copy(int64_t n, const double *__restrict__ s, double *__restrict__ d)
{
    int64_t i = (int64_t)blockIdx.x * OP_THREADS + threadIdx.x;
    if (i < n) d[i] = s[i];
}

__global__ void __launch_bounds__(OP_THREADS) k_fill(int64_t n, double v, double *__restrict__ d)
{
    int64_t i = (int64_t)blockIdx.x * OP_THREADS + threadIdx.x;
    if (i < n) d[i] = v;
}

extern "C" int sph_copy(sph_ctx *ctx, int64_t n, const double *src, double *dst)
{
    if (!ctx) return SPH_EINVAL;
    if (n <= 0) return SPH_OK;
    SPH_REQUIRE(ctx, src && dst, "sph_copy: null");
    hipLaunchKernelGGL(k_copy, dim3(sph_blocks(n, OP_THREADS)), dim3(OP_THREADS), 0, ctx->stream,
                       n, src, dst);
    SPH_LAUNCH_CHECK(ctx, "k_copy");
    return SPH_OK;
}

extern "C" int sph_fill(sph_ctx *ctx, int64_t n, double value, double *dst)
{
    if (!ctx) return SPH_EINVAL;
    if (n <= 0) return SPH_OK;
    SPH_REQUIRE(ctx, dst, "sph_fill: null");
    hipLaunchKernelGGL(k_fill, dim3(sph_blocks(n, OP_THREADS)), dim3(OP_THREADS), 0, ctx->stream,
                       n, value, dst);
    SPH_LAUNCH_CHECK(ctx, "k_fill");
    return SPH_OK;
}

// several copies / fills of n doubles in one launch (save_initial_arrays and
// reset_accelerations of an array's stepped properties)
__global__ void __launch_bounds__(OP_THREADS) k_copy_many(int64_t n, sph_prop_list pl)
{
    const int64_t i = (int64_t)blockIdx.x * OP_THREADS + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < pl.nprops; ++k) pl.dst[k][i] = pl.src[k][i];
}

__global__ void __launch_bounds__(OP_THREADS) k_fill_many(int64_t n, double v, sph_prop_list pl)
{
    const int64_t i = (int64_t)blockIdx.x * OP_THREADS + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < pl.nprops; ++k) pl.dst[k][i] = v;
}

static int check_many(sph_ctx *ctx, const char *name, const sph_prop_list *pl, int need_src)
{
    SPH_REQUIRE(ctx, pl && pl->nprops >= 0 && pl->nprops <= SPH_MAX_PROPS, "%s: bad props",
                name);
    for (int k = 0; k < pl->nprops; ++k)
        SPH_REQUIRE(ctx, pl->dst[k] && (!need_src || pl->src[k]), "%s: null prop %d", name, k);
    return SPH_OK;
}

extern "C" int sph_copy_many(sph_ctx *ctx, int64_t n, const sph_prop_list *props)
{
    if (!ctx) return SPH_EINVAL;
    int rc = check_many(ctx, "sph_copy_many", props, 1);
    if (rc || n <= 0 || props->nprops == 0) return rc;
    hipLaunchKernelGGL(k_copy_many, dim3(sph_blocks(n, OP_THREADS)), dim3(OP_THREADS), 0,
                       ctx->stream, n, *props);
    SPH_LAUNCH_CHECK(ctx, "k_copy_many");
    return SPH_OK;
}

extern "C" int sph_fill_many(sph_ctx *ctx, int64_t n, double value, const sph_prop_list *props)
{
    if (!ctx) return SPH_EINVAL;
    int rc = check_many(ctx, "sph_fill_many", props, 0);
    if (rc || n <= 0 || props->nprops == 0) return rc;
    hipLaunchKernelGGL(k_fill_many, dim3(sph_blocks(n, OP_THREADS)), dim3(OP_THREADS), 0,
                       ctx->stream, n, value, *props);
    SPH_LAUNCH_CHECK(ctx, "k_fill_many");
    return SPH_OK;
}

// The operators beside the path in the reference's dam-break examples:
// the continuity equation (SPHDensityRate), the XSPH velocity correction
// (XSPHCorrection) and the reductions of the viscous / force-based time
// steps.  Same structure as sph_ops.hip: one lane per sorted destination,
// SELL-64 rows, sorted records, reference expression order, no FMA
// contraction.
#include <math.h>
#include <string.h>

#include "sph_internal.h"
#include "sph_terms.h"

#define RX_THREADS 256
#define RX_NT 4  // neighbour records in flight per lane

// ---------------------------------------------------------------------------
// SPHDensityRate (sph/funcs/density_funcs.pyx:116-164; hks and the kernel
// corrections off): rate_a = sum_b (v_a - v_b) . gradW_a(r_ab, h_ab) m_b, in
// list order (SPHFunctionParticle.eval_single, sph_func.pyx:427-441), added
// to the output (SPHFunction.eval, sph_func.pyx:208-227).
// ---------------------------------------------------------------------------
template <bool UH, typename T, typename P>
__device__ __forceinline__ T rate_term(int dim, T xa, T ya, T za, T ua, T va, T wa, T ha, T facA,
                                       P s0, P s1, P s2, P s3, T hb)
{
    T hab, fac;
    if (UH || hb == ha) {
        hab = ha;
        fac = facA;
    } else {
        hab = T(0.5) * (ha + hb);
        fac = sph_h_dim(hab, dim) * sph_cubic_fac1<T>(dim);
    }
    const T rabx = xa - s0.x, raby = ya - s0.y, rabz = za - s1.x;
    const T g = sph_cubic_gfac_fac<T>(rabx, raby, rabz, hab, fac);
    // kernel.gradient: grad = (pa - pb) * (val*fac)
    const T gx = rabx * g, gy = raby * g, gz = rabz * g;
    const T velx = ua - s2.x, vely = va - s2.y, velz = wa - s3.x;
    // cPoint_dot(vel, grad) * m_b (base/point.pxd:23-24)
    return (velx * gx + vely * gy + velz * gz) * s1.y;
}

template <bool UH, typename T>
__global__ void __launch_bounds__(RX_THREADS)
    k_density_rate(int dim, int64_t nd_total, int64_t nd_real, const int32_t *__restrict__ dperm,
                   const int32_t *__restrict__ counts, const int64_t *__restrict__ slice_off,
                   const int32_t *__restrict__ indices, const T *__restrict__ drec,
                   const T *__restrict__ srec, int accumulate, double *__restrict__ out)
{
    const int64_t t = (int64_t)sph_xcd_tile(blockIdx.x, gridDim.x) * RX_THREADS + threadIdx.x;
    if (t >= nd_total) return;
    const int32_t i = dperm[t];
    if (i >= nd_real) {  // non-LocalReal destinations (sph_func.pyx:226-227)
        out[i] = 0;
        return;
    }
    const auto *dr = rec2(drec, t);
    const auto d0 = dr[0], d1 = dr[1], d2 = dr[2], d3 = dr[3];
    const T ha = drec[t * SPH_REC + SPH_R_H];
    const T facA = sph_h_dim(ha, dim) * sph_cubic_fac1<T>(dim);
    T nr = T(0.0);
    const int32_t *row = indices + slice_off[t >> 6] + (t & 63);
    const int32_t e = counts[t];
    int32_t k = 0;
    for (; k + RX_NT <= e; k += RX_NT) {
        int32_t j[RX_NT];
        typename Pair<T>::type s0[RX_NT], s1[RX_NT], s2[RX_NT], s3[RX_NT];
        T hb[RX_NT];
#pragma unroll
        for (int u = 0; u < RX_NT; ++u)
            j[u] = __builtin_nontemporal_load(row + ((int64_t)(k + u) << 6));
#pragma unroll
        for (int u = 0; u < RX_NT; ++u) {
            const auto *sr = rec2(srec, j[u]);
            s0[u] = sr[0];
            s1[u] = sr[1];
            s2[u] = sr[2];
            s3[u] = sr[3];
            hb[u] = UH ? ha : srec[(int64_t)j[u] * SPH_REC + SPH_R_H];
        }
#pragma unroll
        for (int u = 0; u < RX_NT; ++u)
            nr += rate_term<UH, T>(dim, d0.x, d0.y, d1.x, d2.x, d2.y, d3.x, ha, facA, s0[u],
                                   s1[u], s2[u], s3[u], hb[u]);
    }
    for (; k < e; ++k) {
        const int32_t j = __builtin_nontemporal_load(row + ((int64_t)k << 6));
        const auto *sr = rec2(srec, j);
        const T hb = UH ? ha : srec[(int64_t)j * SPH_REC + SPH_R_H];
        nr += rate_term<UH, T>(dim, d0.x, d0.y, d1.x, d2.x, d2.y, d3.x, ha, facA, sr[0], sr[1],
                               sr[2], sr[3], hb);
    }
    out[i] = (accumulate ? T(out[i]) : T(0.0)) + nr;
}

// ---------------------------------------------------------------------------
// XSPHCorrection (sph/funcs/xsph_funcs.pyx:36-91; a CSPHFunctionParticle
// whose denominator stays 0 without the Bonnet-Lok correction, so nothing
// is divided): nr += (m_b W_ab(h_ab) / rho_ab) (v_b - v_a) eps with
// rho_ab = 0.5*(rho_b + rho_a).
// ---------------------------------------------------------------------------
template <bool UH, typename T, typename P>
__device__ __forceinline__ void xsph_term(int dim, T xa, T ya, T za, T ua, T va, T wa, T rhoa,
                                          T ha, T facA, T eps, P s0, P s1, P s2, P s3, P s4,
                                          T hb, T &nr0, T &nr1, T &nr2)
{
    T hab, fac;
    if (UH || hb == ha) {
        hab = ha;
        fac = facA;
    } else {
        hab = T(0.5) * (ha + hb);
        fac = sph_h_dim(hab, dim) * sph_cubic_fac1<T>(dim);
    }
    const T rhoab = T(0.5) * (s4.x + rhoa);
    const T vbax = s2.x - ua, vbay = s2.y - va, vbaz = s3.x - wa;  // cPoint_sub(Vb, Va)
    const T mb = s1.y;
    const T w = sph_cubic_w_fac<T>(xa - s0.x, ya - s0.y, za - s1.x, hab, fac);
    const T temp = mb * w / rhoab;
    nr0 += temp * vbax * eps;
    nr1 += temp * vbay * eps;
    nr2 += temp * vbaz * eps;
}

template <bool UH, typename T>
__global__ void __launch_bounds__(RX_THREADS)
    k_xsph(int dim, int64_t nd_total, int64_t nd_real, const int32_t *__restrict__ dperm,
           const int32_t *__restrict__ counts, const int64_t *__restrict__ slice_off,
           const int32_t *__restrict__ indices, const T *__restrict__ drec,
           const T *__restrict__ srec, T eps, int accumulate, double *__restrict__ ox,
           double *__restrict__ oy, double *__restrict__ oz)
{
    const int64_t t = (int64_t)sph_xcd_tile(blockIdx.x, gridDim.x) * RX_THREADS + threadIdx.x;
    if (t >= nd_total) return;
    const int32_t i = dperm[t];
    if (i >= nd_real) {
        ox[i] = 0;
        if (oy) oy[i] = 0;
        if (oz) oz[i] = 0;
        return;
    }
    const auto *dr = rec2(drec, t);
    const auto d0 = dr[0], d1 = dr[1], d2 = dr[2], d3 = dr[3], d4 = dr[4];
    const T ha = drec[t * SPH_REC + SPH_R_H];
    const T facA = sph_h_dim(ha, dim) * sph_cubic_fac1<T>(dim);
    T nr0 = T(0.0), nr1 = T(0.0), nr2 = T(0.0);
    const int32_t *row = indices + slice_off[t >> 6] + (t & 63);
    const int32_t e = counts[t];
    int32_t k = 0;
    for (; k + RX_NT <= e; k += RX_NT) {
        int32_t j[RX_NT];
        typename Pair<T>::type s0[RX_NT], s1[RX_NT], s2[RX_NT], s3[RX_NT], s4[RX_NT];
        T hb[RX_NT];
#pragma unroll
        for (int u = 0; u < RX_NT; ++u)
            j[u] = __builtin_nontemporal_load(row + ((int64_t)(k + u) << 6));
#pragma unroll
        for (int u = 0; u < RX_NT; ++u) {
            const auto *sr = rec2(srec, j[u]);
            s0[u] = sr[0];
            s1[u] = sr[1];
            s2[u] = sr[2];
            s3[u] = sr[3];
            s4[u] = sr[4];
            hb[u] = UH ? ha : srec[(int64_t)j[u] * SPH_REC + SPH_R_H];
        }
#pragma unroll
        for (int u = 0; u < RX_NT; ++u)
            xsph_term<UH, T>(dim, d0.x, d0.y, d1.x, d2.x, d2.y, d3.x, d4.x, ha, facA, eps, s0[u],
                             s1[u], s2[u], s3[u], s4[u], hb[u], nr0, nr1, nr2);
    }
    for (; k < e; ++k) {
        const int32_t j = __builtin_nontemporal_load(row + ((int64_t)k << 6));
        const auto *sr = rec2(srec, j);
        const T hb = UH ? ha : srec[(int64_t)j * SPH_REC + SPH_R_H];
        xsph_term<UH, T>(dim, d0.x, d0.y, d1.x, d2.x, d2.y, d3.x, d4.x, ha, facA, eps, sr[0],
                         sr[1], sr[2], sr[3], sr[4], hb, nr0, nr1, nr2);
    }
    ox[i] = (accumulate ? T(ox[i]) : T(0.0)) + nr0;
    if (oy) oy[i] = (accumulate ? T(oy[i]) : T(0.0)) + nr1;
    if (oz) oz[i] = (accumulate ? T(oz[i]) : T(0.0)) + nr2;
}

template <typename T>
static int check_list_op(sph_ctx *ctx, const char *name, int dim, int64_t nd_total,
                         int64_t nd_real, const int32_t *d_perm, const int32_t *counts,
                         const int64_t *slice_off, const T *d_rec, const T *s_rec,
                         const double *out)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, dim >= 1 && dim <= 3, "%s: bad dim %d", name, dim);
    SPH_REQUIRE(ctx, nd_real >= 0 && nd_total >= nd_real, "%s: bad n", name);
    if (nd_total == 0) return SPH_OK;
    SPH_REQUIRE(ctx, d_perm && counts && slice_off && d_rec && s_rec && out,
                "%s: null argument", name);
    return SPH_OK;
}

template <typename T>
static int density_rate(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                        const int32_t *d_perm, const int32_t *counts, const int64_t *slice_off,
                        const int32_t *indices, const T *d_rec, const T *s_rec, int uniform_h,
                        int accumulate, double *out)
{
    int rc = check_list_op<T>(ctx, "sph_density_rate", dim, nd_total, nd_real, d_perm, counts,
                              slice_off, d_rec, s_rec, out);
    if (rc || nd_total == 0) return rc;
    auto kern = uniform_h ? k_density_rate<true, T> : k_density_rate<false, T>;
    hipLaunchKernelGGL(kern, dim3(sph_blocks(nd_total, RX_THREADS)), dim3(RX_THREADS), 0,
                       ctx->stream, dim, nd_total, nd_real, d_perm, counts, slice_off, indices,
                       d_rec, s_rec, accumulate, out);
    SPH_LAUNCH_CHECK(ctx, "k_density_rate");
    return SPH_OK;
}

template <typename T>
static int xsph(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real, const int32_t *d_perm,
                const int32_t *counts, const int64_t *slice_off, const int32_t *indices,
                const T *d_rec, const T *s_rec, int uniform_h, double eps, int accumulate,
                double *ox, double *oy, double *oz)
{
    int rc = check_list_op<T>(ctx, "sph_xsph", dim, nd_total, nd_real, d_perm, counts,
                              slice_off, d_rec, s_rec, ox);
    if (rc || nd_total == 0) return rc;
    auto kern = uniform_h ? k_xsph<true, T> : k_xsph<false, T>;
    hipLaunchKernelGGL(kern, dim3(sph_blocks(nd_total, RX_THREADS)), dim3(RX_THREADS), 0,
                       ctx->stream, dim, nd_total, nd_real, d_perm, counts, slice_off, indices,
                       d_rec, s_rec, T(eps), accumulate, ox, oy, oz);
    SPH_LAUNCH_CHECK(ctx, "k_xsph");
    return SPH_OK;
}

extern "C" int sph_density_rate(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                                const int32_t *d_perm, const int32_t *counts,
                                const int64_t *slice_off, const int32_t *indices,
                                const double *d_rec, const double *s_rec, int uniform_h,
                                int accumulate, double *out)
{
    return density_rate<double>(ctx, dim, nd_total, nd_real, d_perm, counts, slice_off, indices,
                                d_rec, s_rec, uniform_h, accumulate, out);
}

extern "C" int sph_density_rate_f32(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                                    const int32_t *d_perm, const int32_t *counts,
                                    const int64_t *slice_off, const int32_t *indices,
                                    const float *d_rec, const float *s_rec, int uniform_h,
                                    int accumulate, double *out)
{
    return density_rate<float>(ctx, dim, nd_total, nd_real, d_perm, counts, slice_off, indices,
                               d_rec, s_rec, uniform_h, accumulate, out);
}

extern "C" int sph_xsph(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                        const int32_t *d_perm, const int32_t *counts, const int64_t *slice_off,
                        const int32_t *indices, const double *d_rec, const double *s_rec,
                        int uniform_h, double eps, int accumulate, double *ox, double *oy,
                        double *oz)
{
    return xsph<double>(ctx, dim, nd_total, nd_real, d_perm, counts, slice_off, indices, d_rec,
                        s_rec, uniform_h, eps, accumulate, ox, oy, oz);
}

extern "C" int sph_xsph_f32(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                            const int32_t *d_perm, const int32_t *counts,
                            const int64_t *slice_off, const int32_t *indices,
                            const float *d_rec, const float *s_rec, int uniform_h, double eps,
                            int accumulate, double *ox, double *oy, double *oz)
{
    return xsph<float>(ctx, dim, nd_total, nd_real, d_perm, counts, slice_off, indices, d_rec,
                       s_rec, uniform_h, eps, accumulate, ox, oy, oz);
}

// ---------------------------------------------------------------------------
// reductions of the time-step functions
// ---------------------------------------------------------------------------
// NaN-propagating max and min, like numpy.max / numpy.min
__device__ __forceinline__ double nmax(double a, double b)
{
    if (a != a) return a;
    if (b != b) return b;
    return b > a ? b : a;
}
__device__ __forceinline__ double nmin(double a, double b)
{
    if (a != a) return a;
    if (b != b) return b;
    return b < a ? b : a;
}

// per block: max_i |a_i| over ncomp components (sqrt(x*x + y*y [+ z*z]) in
// numpy's order; |x| for one component), min and max of v
#define RED_BLOCKS 512

__device__ __forceinline__ double wave_nmax(double v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = nmax(v, __shfl_xor(v, d, 64));
    return v;
}
__device__ __forceinline__ double wave_nmin(double v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = nmin(v, __shfl_xor(v, d, 64));
    return v;
}

__global__ void __launch_bounds__(RX_THREADS)
    k_extrema_part(int64_t n, int ncomp, const double *__restrict__ a,
                   const double *__restrict__ b, const double *__restrict__ c,
                   double *__restrict__ part)
{
    __shared__ double s_mx[RX_THREADS / 64], s_mn[RX_THREADS / 64], s_mm[RX_THREADS / 64];
    double mx = -INFINITY, mn = INFINITY, mm = -INFINITY;
    for (int64_t i = (int64_t)blockIdx.x * RX_THREADS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * RX_THREADS) {
        const double x = a[i];
        double v;
        if (ncomp == 1) {
            v = fabs(x);
        } else {
            const double y = b[i];
            double s = x * x + y * y;
            if (ncomp == 3) {
                const double z = c[i];
                s = s + z * z;
            }
            v = sqrt(s);
        }
        mx = nmax(mx, v);   // |a| over the components
        mn = nmin(mn, x);   // min and max of the first array alone
        mm = nmax(mm, x);
    }
    mx = wave_nmax(mx);
    mn = wave_nmin(mn);
    mm = wave_nmax(mm);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        s_mx[wid] = mx;
        s_mn[wid] = mn;
        s_mm[wid] = mm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < RX_THREADS / 64; ++w) {
            mx = nmax(s_mx[0], s_mx[w]);
            s_mx[0] = mx;
            s_mn[0] = nmin(s_mn[0], s_mn[w]);
            s_mm[0] = nmax(s_mm[0], s_mm[w]);
        }
        part[3 * blockIdx.x] = s_mx[0];
        part[3 * blockIdx.x + 1] = s_mn[0];
        part[3 * blockIdx.x + 2] = s_mm[0];
    }
}

__global__ void __launch_bounds__(64) k_extrema_final(int nb, const double *__restrict__ part,
                                                      double *__restrict__ out)
{
    double mx = -INFINITY, mn = INFINITY, mm = -INFINITY;
    for (int k = threadIdx.x; k < nb; k += 64) {
        mx = nmax(mx, part[3 * k]);
        mn = nmin(mn, part[3 * k + 1]);
        mm = nmax(mm, part[3 * k + 2]);
    }
    mx = wave_nmax(mx);
    mn = wave_nmin(mn);
    mm = wave_nmax(mm);
    if (threadIdx.x == 0) {
        out[0] = mx;
        out[1] = mn;
        out[2] = mm;
    }
}

extern "C" int sph_extrema(sph_ctx *ctx, int64_t n, int ncomp, const double *a, const double *b,
                           const double *c, double *out)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, out && n >= 0 && ncomp >= 1 && ncomp <= 3, "sph_extrema: bad argument");
    SPH_REQUIRE(ctx, n == 0 || (a && (ncomp < 2 || b) && (ncomp < 3 || c)),
                "sph_extrema: null array");
    if (n == 0) {
        out[0] = -INFINITY;
        out[1] = INFINITY;
        out[2] = -INFINITY;
        return SPH_OK;
    }
    int64_t nb = (n + RX_THREADS - 1) / RX_THREADS;
    if (nb > RED_BLOCKS) nb = RED_BLOCKS;
    int rc = sph_ensure_scratch(ctx, (size_t)(3 * nb + 4) * sizeof(double));
    if (rc) return rc;
    rc = sph_ensure_pinned(ctx, 4 * sizeof(double));
    if (rc) return rc;
    double *part = (double *)ctx->scratch;
    double *res = part + 3 * nb;
    hipLaunchKernelGGL(k_extrema_part, dim3((unsigned)nb), dim3(RX_THREADS), 0, ctx->stream, n,
                       ncomp, a, b, c, part);
    SPH_LAUNCH_CHECK(ctx, "k_extrema_part");
    hipLaunchKernelGGL(k_extrema_final, dim3(1), dim3(64), 0, ctx->stream, (int)nb,
                       (const double *)part, res);
    SPH_LAUNCH_CHECK(ctx, "k_extrema_final");
    SPH_HIP_CHECK(ctx, hipMemcpyAsync(ctx->pinned, res, 3 * sizeof(double),
                                      hipMemcpyDeviceToHost, ctx->stream));
    SPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    memcpy(out, ctx->pinned, 3 * sizeof(double));
    return SPH_OK;
}

// ---------------------------------------------------------------------------
// the same reductions without allocating: partials in a file-scope device
// array (EXTREMA_SCRATCH_DOUBLES doubles, 12 KiB per device, held until the
// process ends) and the three results written to device memory
// ---------------------------------------------------------------------------
#define EXTREMA_SCRATCH_DOUBLES (3 * RED_BLOCKS)
static __device__ __align__(256) double scratch[EXTREMA_SCRATCH_DOUBLES];

// The device address of `scratch` on the context's device (each device holds
// its own copy).  Called by sph_ctx_create with the device current (and by
// sph_extrema_noalloc if that failed); host code uses only this address, never
// the array's name.
int sph_extrema_scratch_setup(sph_ctx *ctx)
{
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return sph_set_error(ctx, SPH_EHIP, "hipGetSymbolAddress(extrema scratch) failed: %s",
                             hipGetErrorString(e));
    ctx->extrema_scratch = (double *)p;
    return SPH_OK;
}

extern "C" int sph_extrema_noalloc(sph_ctx *ctx, int64_t n, int ncomp, const double *a,
                                   const double *b, const double *c, double *out_dev)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, out_dev && n >= 0 && ncomp >= 1 && ncomp <= 3,
                "sph_extrema_noalloc: bad argument");
    SPH_REQUIRE(ctx, n == 0 || (a && (ncomp < 2 || b) && (ncomp < 3 || c)),
                "sph_extrema_noalloc: null array");
    if (!ctx->extrema_scratch) {
        int cur = 0;
        SPH_HIP_CHECK(ctx, hipGetDevice(&cur));
        SPH_HIP_CHECK(ctx, hipSetDevice(ctx->device));
        const int rc = sph_extrema_scratch_setup(ctx);
        SPH_HIP_CHECK(ctx, hipSetDevice(cur));
        if (rc) return rc;
    }
    // n == 0: the final pass over no partials writes -inf, +inf, -inf
    int64_t nb = (n + RX_THREADS - 1) / RX_THREADS;
    if (nb > RED_BLOCKS) nb = RED_BLOCKS;
    SPH_REQUIRE(ctx, 3 * nb <= EXTREMA_SCRATCH_DOUBLES,
                "sph_extrema_noalloc: %lld partials exceed the scratch of %d doubles",
                (long long)(3 * nb), EXTREMA_SCRATCH_DOUBLES);
    double *part = ctx->extrema_scratch;
    if (nb > 0) {
        hipLaunchKernelGGL(k_extrema_part, dim3((unsigned)nb), dim3(RX_THREADS), 0, ctx->stream,
                           n, ncomp, a, b, c, part);
        SPH_LAUNCH_CHECK(ctx, "k_extrema_part");
    }
    hipLaunchKernelGGL(k_extrema_final, dim3(1), dim3(64), 0, ctx->stream, (int)nb,
                       (const double *)part, out_dev);
    SPH_LAUNCH_CHECK(ctx, "k_extrema_final");
    return SPH_OK;
}

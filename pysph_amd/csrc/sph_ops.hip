// SPH operators over neighbour lists and the integrator stage arithmetic.
//
// Every per-pair expression follows the reference Cython in order
// (sph/funcs/density_funcs.pyx:37-73, sph/funcs/eos_funcs.pyx:118-129,
// sph/funcs/pressure_funcs.pyx:246-391) and is compiled without FMA
// contraction, so with the lists in the reference's order the sums are the
// reference's sums.
#include "sph_internal.h"

#define OP_THREADS 256

// ---------------------------------------------------------------------------
// SPHRho
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(OP_THREADS)
    k_summation_density(int dim, int64_t nd_real, int64_t nd_total,
                        const int64_t *__restrict__ offsets, const int32_t *__restrict__ indices,
                        const double *__restrict__ dx, const double *__restrict__ dy,
                        const double *__restrict__ dz, const double *__restrict__ dh,
                        const double *__restrict__ sx, const double *__restrict__ sy,
                        const double *__restrict__ sz, const double *__restrict__ sh,
                        const double *__restrict__ sm, int accumulate, double *__restrict__ rho)
{
    int64_t i = (int64_t)blockIdx.x * OP_THREADS + threadIdx.x;
    if (i >= nd_total) return;
    if (i >= nd_real) {  // non-LocalReal destinations (sph_func.pyx:226-227)
        rho[i] = 0;
        return;
    }
    const double xa = dx[i], ya = dy[i], za = dz[i], ha = dh[i];
    double nr = 0.0;
    const int64_t e = offsets[i + 1];
    for (int64_t k = offsets[i]; k < e; ++k) {
        const int32_t j = indices[k];
        const double hb = sh[j];
        const double hab = 0.5 * (ha + hb);
        const double w = sph_cubic_w(dim, xa - sx[j], ya - sy[j], za - sz[j], hab);
        nr += w * sm[j];
    }
    rho[i] = (accumulate ? rho[i] : 0.0) + nr;
}

extern "C" int sph_summation_density(sph_ctx *ctx, int dim, int64_t nd_real, int64_t nd_total,
                                     const int64_t *offsets, const int32_t *indices,
                                     const double *dx, const double *dy, const double *dz,
                                     const double *dh, const double *sx, const double *sy,
                                     const double *sz, const double *sh, const double *sm,
                                     int accumulate, double *rho)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, dim >= 1 && dim <= 3, "sph_summation_density: bad dim %d", dim);
    SPH_REQUIRE(ctx, nd_real >= 0 && nd_total >= nd_real, "sph_summation_density: bad n");
    if (nd_total == 0) return SPH_OK;
    SPH_REQUIRE(ctx, rho && (nd_real == 0 || (offsets && dx && dy && dz && dh)),
                "sph_summation_density: null argument");
    hipLaunchKernelGGL(k_summation_density, dim3(sph_blocks(nd_total, OP_THREADS)),
                       dim3(OP_THREADS), 0, ctx->stream, dim, nd_real, nd_total, offsets,
                       indices, dx, dy, dz, dh, sx, sy, sz, sh, sm, accumulate, rho);
    SPH_LAUNCH_CHECK(ctx, "k_summation_density");
    return SPH_OK;
}

// ---------------------------------------------------------------------------
// Tait equation of state
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(OP_THREADS)
    k_tait(int64_t n_real, int64_t n_total, const double *__restrict__ rho, double ro,
           double co, double gamma, double B, int accumulate, double *__restrict__ p,
           double *__restrict__ cs)
{
    int64_t i = (int64_t)blockIdx.x * OP_THREADS + threadIdx.x;
    if (i >= n_total) return;
    if (i >= n_real) {
        p[i] = 0;
        cs[i] = 0;
        return;
    }
    const double rhoa = rho[i];
    const double ratio = rhoa / ro;
    const double gamma2 = 0.5 * (gamma - 1.0);
    const double tmp = pow(ratio, gamma);
    const double r0 = (tmp - 1.0) * B;
    const double r1 = pow(ratio, gamma2) * co;
    p[i] = (accumulate ? p[i] : 0.0) + r0;
    cs[i] = (accumulate ? cs[i] : 0.0) + r1;
}

extern "C" int sph_tait(sph_ctx *ctx, int64_t n_real, int64_t n_total, const double *rho,
                        double ro, double co, double gamma, int accumulate, double *p,
                        double *cs)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, n_real >= 0 && n_total >= n_real, "sph_tait: bad n");
    if (n_total == 0) return SPH_OK;
    SPH_REQUIRE(ctx, rho && p && cs, "sph_tait: null argument");
    const double B = co * co * ro / gamma;  // eos_funcs.pyx:90
    hipLaunchKernelGGL(k_tait, dim3(sph_blocks(n_total, OP_THREADS)), dim3(OP_THREADS), 0,
                       ctx->stream, n_real, n_total, rho, ro, co, gamma, B, accumulate, p, cs);
    SPH_LAUNCH_CHECK(ctx, "k_tait");
    return SPH_OK;
}

// ---------------------------------------------------------------------------
// MomentumEquation with Monaghan artificial viscosity
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(OP_THREADS)
    k_momentum(int dim, int64_t nd_real, int64_t nd_total, const int64_t *__restrict__ offsets,
               const int32_t *__restrict__ indices, sph_mom_dst d, sph_mom_src s, double alpha,
               double beta, double eta, int accumulate, double *__restrict__ ax,
               double *__restrict__ ay, double *__restrict__ az, double *__restrict__ dt_fac)
{
    int64_t i = (int64_t)blockIdx.x * OP_THREADS + threadIdx.x;
    if (i >= nd_total) return;
    if (i >= nd_real) {
        ax[i] = 0;
        ay[i] = 0;
        if (az) az[i] = 0;
        return;
    }
    const double xa = d.x[i], ya = d.y[i], za = d.z[i], ha = d.h[i];
    const double ua = d.u[i], va = d.v[i], wa = d.w[i];
    const double ca = d.cs[i], Pa = d.p[i], rhoa = d.rho[i];
    const double rhoa2 = 1.0 / (rhoa * rhoa);
    double dtf = dt_fac[i];
    double nr0 = 0.0, nr1 = 0.0, nr2 = 0.0;
    const int64_t e = offsets[i + 1];
    for (int64_t k = offsets[i]; k < e; ++k) {
        const int32_t j = indices[k];
        const double hb = s.h[j];
        const double hab = 0.5 * (ha + hb);
        const double cb = s.cs[j];
        const double rabx = xa - s.x[j];
        const double raby = ya - s.y[j];
        const double rabz = za - s.z[j];
        const double vabx = ua - s.u[j];
        const double vaby = va - s.v[j];
        const double vabz = wa - s.w[j];
        const double dot = vabx * rabx + vaby * raby + vabz * rabz;
        const double rab2 = rabx * rabx + raby * raby + rabz * rabz;
        const double den = rab2 + eta * eta * hab * hab;
        const double q = hab * dot / den;  // the same quotient as mu below
        const double f = fabs(q);
        if (f > dtf) dtf = f;
        const double Pb = s.p[j], rhob = s.rho[j], mb = s.m[j];
        const double rhob2 = 1.0 / (rhob * rhob);
        double tmp = Pa * rhoa2 + Pb * rhob2;
        double piab = 0;
        if (dot < 0) {
            const double cab = 0.5 * (ca + cb);
            const double rhoab = 0.5 * (rhoa + rhob);
            const double mu = q;  // mu = hab*dot; mu /= (rab2 + eta*eta*hab*hab)
            piab = -alpha * cab * mu + beta * mu * mu;
            piab /= rhoab;
        }
        tmp += piab;
        tmp *= -mb;
        const double g = sph_cubic_gfac(dim, rabx, raby, rabz, hab);
        nr0 += tmp * (rabx * g);
        nr1 += tmp * (raby * g);
        nr2 += tmp * (rabz * g);
    }
    dt_fac[i] = dtf;
    ax[i] = (accumulate ? ax[i] : 0.0) + nr0;
    ay[i] = (accumulate ? ay[i] : 0.0) + nr1;
    if (az) az[i] = (accumulate ? az[i] : 0.0) + nr2;
}

extern "C" int sph_momentum_av(sph_ctx *ctx, int dim, int64_t nd_real, int64_t nd_total,
                               const int64_t *offsets, const int32_t *indices,
                               const sph_mom_dst *dst, const sph_mom_src *src, double alpha,
                               double beta, double eta, int accumulate, double *ax, double *ay,
                               double *az, double *dt_fac)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, dim >= 1 && dim <= 3, "sph_momentum_av: bad dim %d", dim);
    SPH_REQUIRE(ctx, nd_real >= 0 && nd_total >= nd_real, "sph_momentum_av: bad n");
    if (nd_total == 0) return SPH_OK;
    SPH_REQUIRE(ctx, dst && src && ax && ay && dt_fac, "sph_momentum_av: null argument");
    hipLaunchKernelGGL(k_momentum, dim3(sph_blocks(nd_total, OP_THREADS)), dim3(OP_THREADS), 0,
                       ctx->stream, dim, nd_real, nd_total, offsets, indices, *dst, *src, alpha,
                       beta, eta, accumulate, ax, ay, az, dt_fac);
    SPH_LAUNCH_CHECK(ctx, "k_momentum");
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
    SPH_REQUIRE(ctx, props && props->nprops >= 0 && props->nprops <= SPH_MAX_STEP_PROPS,
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
    SPH_REQUIRE(ctx, props && props->nprops >= 0 && props->nprops <= SPH_MAX_STEP_PROPS,
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

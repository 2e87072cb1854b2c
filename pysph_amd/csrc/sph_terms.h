// Per-pair terms of the neighbour operators, shared by the operator kernels
// and the fused kernels of the batched path.  Every expression follows the
// reference Cython in order (sph/funcs/density_funcs.pyx:37-73,
// sph/funcs/pressure_funcs.pyx:246-391); the sources are compiled without
// FMA contraction.
#pragma once
#include "sph_internal.h"

// Sorted records: fields SPH_R_* of SPH_REC values of the arithmetic type T
// (double, or float for the single-precision variant), 16-byte (8-byte)
// aligned; the neighbour loops fetch them as pairs.
template <typename T> struct Pair;
template <> struct Pair<double> { using type = double2; };
template <> struct Pair<float> { using type = float2; };

template <typename T>
__device__ __forceinline__ const typename Pair<T>::type *rec2(const T *rec, int64_t j)
{
    return reinterpret_cast<const typename Pair<T>::type *>(rec + j * SPH_REC);
}

// ---------------------------------------------------------------------------
// SPHRho
// ---------------------------------------------------------------------------
// one neighbour term of SPHRho (density_funcs.pyx:41-73); sxy = (x, y) and
// szm = (z, m) of the source record, hb its h (not read when UH)
// invA: the destination's sph_inv_for_div(h_a), so q = r/h_a is the IEEE
// quotient from three operations (sph_div_inv)
template <bool UH, typename T, typename P>
__device__ __forceinline__ T rho_term(int dim, T xa, T ya, T za, T ha, T facA, P sxy, P szm,
                                      T hb, T invA)
{
    // with h_b == h_a, 0.5*(h_a+h_b) == h_a exactly and h_dim(h)*sigma is the
    // destination's value: reuse it instead of dividing again
    if (UH || hb == ha)
        return sph_cubic_w_fac<T>(xa - sxy.x, ya - sxy.y, za - szm.x, ha, facA, invA) * szm.y;
    const T hab = T(0.5) * (ha + hb);
    const T fac = sph_h_dim(hab, dim) * sph_cubic_fac1<T>(dim);
    return sph_cubic_w_fac<T>(xa - sxy.x, ya - sxy.y, za - szm.x, hab, fac) * szm.y;
}

// ---------------------------------------------------------------------------
// MomentumEquation with Monaghan artificial viscosity
// ---------------------------------------------------------------------------
template <typename T> struct MomDest {
    T xa, ya, za, ha, facA, rhoa, ca, Pa, ua, va, wa, rhoa2;
    T inv_ha;  // sph_inv_for_div(ha)
};

// one neighbour term of MomentumEquation (pressure_funcs.pyx:252-391).
// Source record slots: s0 = (x, y), s1 = (z, m), s2 = (u, v), s3 = (w, P/rho^2),
// s4 = (rho, cs); hb its h (not read when UH).
template <bool UH, typename T, typename P>
__device__ __forceinline__ void mom_term(int dim, const MomDest<T> &D, T alpha, T beta, T eta,
                                         const P s0, const P s1, const P s2, const P s3,
                                         const P s4, T hb, T &dtf, T &nr0, T &nr1, T &nr2)
{
    T hab, fac, inv;
    if (UH || hb == D.ha) {  // 0.5*(h_a+h_a) == h_a: reuse the destination's h_dim(h)*sigma
        hab = D.ha;
        fac = D.facA;
        inv = D.inv_ha;
    } else {
        hab = T(0.5) * (D.ha + hb);
        fac = sph_h_dim(hab, dim) * sph_cubic_fac1<T>(dim);
        inv = T(0.0);
    }
    const T rabx = D.xa - s0.x;
    const T raby = D.ya - s0.y;
    const T rabz = D.za - s1.x;
    const T vabx = D.ua - s2.x;
    const T vaby = D.va - s2.y;
    const T vabz = D.wa - s3.x;
    const T dot = vabx * rabx + vaby * raby + vabz * rabz;
    const T rab2 = rabx * rabx + raby * raby + rabz * rabz;
    const T den = rab2 + eta * eta * hab * hab;
    const T q = hab * dot / den;  // the same quotient as mu below
    const T f = fabs(q);
    dtf = f > dtf ? f : dtf;
    const T mb = s1.y, rhob = s4.x, cb = s4.y;
    // Pb * (1.0/(rhob*rhob)), computed once per source particle (k_prho2)
    T tmp = D.Pa * D.rhoa2 + s3.y;
    // Monaghan viscosity, applied when dot < 0; evaluated for every pair and
    // selected, which costs less than a divergent branch in a wavefront
    const T cab = T(0.5) * (D.ca + cb);
    const T rhoab = T(0.5) * (D.rhoa + rhob);
    const T mu = q;  // mu = hab*dot; mu /= (rab2 + eta*eta*hab*hab)
    T pi_v = -alpha * cab * mu + beta * mu * mu;
    pi_v /= rhoab;
    const T piab = dot < 0 ? pi_v : T(0.0);
    tmp += piab;
    tmp *= -mb;
    const T g = sph_cubic_gfac_fac<T>(rabx, raby, rabz, hab, fac, inv);
    nr0 += tmp * (rabx * g);
    nr1 += tmp * (raby * g);
    nr2 += tmp * (rabz * g);
}

// ---------------------------------------------------------------------------
// Sources of one destination: every function of a calc over its own source
// array (SPHCalc.sph, sph/sph_calc.pyx:307-358) in one launch.  Each source's
// row is summed from 0 in list order and added to the output in source
// order, exactly as separate calls would (out = (out + r_1) + r_2 ...).
// ---------------------------------------------------------------------------
template <typename T> struct SrcLists {
    int nsrc;
    const int32_t *counts[SPH_MAX_SRC];
    const int64_t *slice_off[SPH_MAX_SRC];
    const int32_t *indices[SPH_MAX_SRC];
    const T *srec[SPH_MAX_SRC];
    const T *spm[SPH_MAX_SRC];  // compact (x, y, z, m) rows, or NULL: the records
};

// pairs 0 and 1 of source j ((x, y) and (z, m)): from the compact rows when
// given (4 values per particle), else from the records
template <typename T>
__device__ __forceinline__ const typename Pair<T>::type *pm2(const T *spm, const T *srec,
                                                             int64_t j)
{
    return spm ? reinterpret_cast<const typename Pair<T>::type *>(spm + j * 4) : rec2(srec, j);
}

template <typename T>
static int src_lists(sph_ctx *ctx, const char *name, const sph_sources *src, SrcLists<T> *L)
{
    SPH_REQUIRE(ctx, src && src->nsrc >= 1 && src->nsrc <= SPH_MAX_SRC, "%s: bad nsrc", name);
    L->nsrc = src->nsrc;
    for (int s = 0; s < src->nsrc; ++s) {
        SPH_REQUIRE(ctx, src->counts[s] && src->slice_off[s] && src->s_rec[s] && src->ns[s] >= 0,
                    "%s: null source %d", name, s);
        L->counts[s] = src->counts[s];
        L->slice_off[s] = src->slice_off[s];
        L->indices[s] = src->indices[s];
        L->srec[s] = (const T *)src->s_rec[s];
        L->spm[s] = (const T *)src->s_pm[s];
    }
    return SPH_OK;
}

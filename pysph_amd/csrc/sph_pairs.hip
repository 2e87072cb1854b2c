// Split-form pair operators and the wall force, over the neighbour lists and
// sorted records of the momentum equation:
//   SPHPressureGradient          sph/funcs/pressure_funcs.pyx:19-168
//   MonaghanArtificialViscosity  sph/funcs/viscosity_funcs.pyx:19-157
//   MomentumEquation, deltap     sph/funcs/pressure_funcs.pyx:246-391, :327-365
//   MonaghanBoundaryForce        sph/funcs/boundary_funcs.pyx:9-107
// Every per-pair expression follows the reference Cython in order and is
// compiled without FMA contraction.  The reference modules are compiled
// without cdivision: a float division by zero raises inside a cdef void
// eval_nbr, Cython reports and ignores it, and the pair adds nothing from
// that point on.  The divisions that can meet a zero in use (W(deltap), the
// wall distance y, delp, m_a + m_b, h) end the pair the same way here.
#include "sph_internal.h"
#include "sph_terms.h"

#define PAIR_THREADS 256
// neighbour records whose loads are in flight before their terms are added
#define PAIR_NT 4

namespace {

template <typename T> struct PairDest {
    T xa, ya, za, ha, facA, wdpA, ma, rhoa, rhoa2, ca, Pa, ua, va, wa;
};

template <typename T> struct PairArgs {
    T alpha, beta, eta, deltap, n, eps, delp;
};

struct Frames {
    const int32_t *perm[SPH_MAX_SRC];
    const double *f[SPH_MAX_SRC][6];  // tx ty tz nx ny nz
};

// record pairs each operator reads: (x,y) (z,m) (u,v) (w,P/rho^2) (rho,cs) (h,p)
template <int OP> __host__ __device__ constexpr unsigned need_mask()
{
    return OP == SPH_PAIR_PRESSURE_GRADIENT    ? 0x3bu  // 0 1 3 4 5
           : OP == SPH_PAIR_ARTIFICIAL_VISCOSITY ? 0x3fu
           : OP == SPH_PAIR_MOMENTUM             ? 0x3fu
                                                 : 0x03u;  // boundary force: 0 1
}

// h_ab and h_dim(h_ab)*sigma; with h_b == h_a the destination's values
template <bool UH, typename T>
__device__ __forceinline__ void pair_h(int dim, const PairDest<T> &D, T hb, T &hab, T &fac)
{
    if (UH || hb == D.ha) {
        hab = D.ha;
        fac = D.facA;
    } else {
        hab = T(0.5) * (D.ha + hb);
        fac = sph_h_dim(hab, dim) * sph_cubic_fac1<T>(dim);
    }
}

// artificial pressure factor (W(r_ab)/W(deltap))^n, pressure_funcs.pyx:121-136
// and :350-362 (hks off); false when W(deltap) == 0 (the division raises)
template <bool UH, typename T>
__device__ __forceinline__ bool art_fab(const PairDest<T> &D, const PairArgs<T> &A, T rx, T ry,
                                        T rz, T hb, T hab, T fac, T &fab)
{
    const T w = sph_cubic_w_fac<T>(rx, ry, rz, hab, fac);
    const T wdeltap = (UH || hb == D.ha)
                          ? D.wdpA
                          : sph_cubic_w_fac<T>(A.deltap, T(0.0), T(0.0), hab, fac);
    if (wdeltap == T(0.0)) return false;
    fab = w / wdeltap;
    fab = sph_pow(fab, A.n);
    return true;
}

template <int OP, bool AP, bool UH, typename T, typename P>
__device__ __forceinline__ void pair_term(int dim, const PairDest<T> &D, const PairArgs<T> &A,
                                          const P *s, int32_t o, const double *const *fr,
                                          T &dtf, T &nr0, T &nr1, T &nr2)
{
    const T rabx = D.xa - s[0].x;
    const T raby = D.ya - s[0].y;
    const T rabz = D.za - s[1].x;
    const T mb = s[1].y;
    if constexpr (OP == SPH_PAIR_BOUNDARY_FORCE) {
        // boundary_funcs.pyx:60-107: h, m_a and cs are the destination's
        const T tx = T(fr[0][o]), ty = T(fr[1][o]), tz = T(fr[2][o]);
        const T nx = T(fr[3][o]), ny = T(fr[4][o]), nz = T(fr[5][o]);
        const T h = D.ha, ma = D.ma, cs = D.ca;
        const T x = rabx * tx + raby * ty + rabz * tz;
        const T y = rabx * nx + raby * ny + rabz * nz;
        T force = T(0.0);
        if (h == T(0.0)) return;
        const T q = y / h;
        const T ax = fabs(x);
        if (T(0) <= ax && ax <= A.delp) {
            if (y == T(0.0)) return;
            const T beta = T(0.02) * cs * cs / y;
            if (A.delp == T(0.0)) return;
            const T tforce = T(1.0) - ax / A.delp;
            const T c23 = T(2.0 / 3.0);
            T nforce;
            if (T(0) < q && q <= c23)
                nforce = c23;
            else if (c23 < q && q <= T(1.0))
                nforce = T(2) * q * (T(1.0) - T(0.75) * q);
            else if (T(1.0) < q && q <= T(2.0))
                nforce = T(0.5) * (T(2) - q) * (T(2) - q);
            else
                nforce = T(0.0);
            if (ma + mb == T(0.0)) return;
            force = (mb / (ma + mb)) * nforce * tforce * beta;
        }
        nr0 += force * nx;
        nr1 += force * ny;
        nr2 += force * nz;
    } else {
        const T hb = UH ? D.ha : s[5].x;
        T hab, fac;
        pair_h<UH, T>(dim, D, hb, hab, fac);
        T tmp;
        if constexpr (OP == SPH_PAIR_PRESSURE_GRADIENT) {
            // pressure_funcs.pyx:90-151; s[3].y = Pb * (1.0/(rhob*rhob))
            tmp = -mb * (D.Pa * D.rhoa2 + s[3].y);
            T artificial_stress = T(0.0);
            if constexpr (AP) {
                const T pb = s[5].y, rhob = s[4].x;
                const T rhob2 = T(1.0) / (rhob * rhob);
                T Ra = D.Pa < 0 ? -A.eps * D.Pa : T(0.0);
                Ra = Ra * D.rhoa2;
                T Rb = pb < 0 ? -A.eps * pb : T(0.0);
                Rb = Rb * rhob2;
                T fab;
                if (!art_fab<UH, T>(D, A, rabx, raby, rabz, hb, hab, fac, fab)) return;
                artificial_stress = mb * (Ra + Rb) * fab;
            }
            tmp = tmp + artificial_stress;
        } else {
            const T vabx = D.ua - s[2].x;
            const T vaby = D.va - s[2].y;
            const T vabz = D.wa - s[3].x;
            const T dot = vabx * rabx + vaby * raby + vabz * rabz;
            const T rab2 = rabx * rabx + raby * raby + rabz * rabz;
            const T rhob = s[4].x, cb = s[4].y;
            const T cab = T(0.5) * (D.ca + cb);
            const T rhoab = T(0.5) * (D.rhoa + rhob);
            const T den = rab2 + A.eta * A.eta * hab * hab;
            if constexpr (OP == SPH_PAIR_ARTIFICIAL_VISCOSITY) {
                // viscosity_funcs.pyx:97-131: dt_fac with 0.01 h^2, mu with eta
                const T f = fabs(hab * dot / (rab2 + T(0.01) * hab * hab));
                dtf = f > dtf ? f : dtf;
                if (dot < 0 && den == T(0.0)) return;
                T mu = hab * dot;
                mu /= den;
                T pi_v = -A.alpha * cab * mu + A.beta * mu * mu;
                pi_v /= rhoab;
                tmp = dot < 0 ? pi_v : T(0.0);
                tmp *= -mb;
            } else {
                // pressure_funcs.pyx:283-365
                const T q = hab * dot / den;
                const T f = fabs(q);
                dtf = f > dtf ? f : dtf;
                tmp = D.Pa * D.rhoa2 + s[3].y;
                const T mu = q;
                T pi_v = -A.alpha * cab * mu + A.beta * mu * mu;
                pi_v /= rhoab;
                tmp += dot < 0 ? pi_v : T(0.0);
                tmp *= -mb;
                if constexpr (AP) {
                    const T pb = s[5].y;
                    const T rhob2 = T(1.0) / (rhob * rhob);
                    T Ra = T(0.0);
                    if (D.Pa < 0) Ra = -A.eps * D.Pa;
                    Ra = Ra * D.rhoa2;
                    T Rb = T(0.0);
                    if (pb < 0) Rb = -A.eps * pb;
                    Rb = Rb * rhob2;
                    T fab;
                    if (!art_fab<UH, T>(D, A, rabx, raby, rabz, hb, hab, fac, fab)) return;
                    tmp = tmp + mb * (Ra + Rb) * fab;
                }
            }
        }
        const T g = sph_cubic_gfac_fac<T>(rabx, raby, rabz, hab, fac);
        nr0 += tmp * (rabx * g);
        nr1 += tmp * (raby * g);
        nr2 += tmp * (rabz * g);
    }
}

// one SELL-64 row, summed from 0 in list order
template <int OP, bool AP, bool UH, typename T>
__device__ __forceinline__ void pair_row(int dim, const PairDest<T> &D, const PairArgs<T> &A,
                                         const int32_t *__restrict__ row, int32_t e,
                                         const T *__restrict__ srec,
                                         const int32_t *__restrict__ fperm,
                                         const double *const *fr, T &dtf, T &nr0, T &nr1,
                                         T &nr2)
{
    using P = typename Pair<T>::type;
    constexpr unsigned need = need_mask<OP>();
    constexpr bool frames = OP == SPH_PAIR_BOUNDARY_FORCE;
    int32_t k = 0;
    for (; k + PAIR_NT <= e; k += PAIR_NT) {
        int32_t j[PAIR_NT], o[PAIR_NT];
        P s[PAIR_NT][6];
#pragma unroll
        for (int u = 0; u < PAIR_NT; ++u)
            j[u] = __builtin_nontemporal_load(row + ((int64_t)(k + u) << 6));
#pragma unroll
        for (int u = 0; u < PAIR_NT; ++u) {
            const P *sr = rec2(srec, j[u]);
#pragma unroll
            for (int f = 0; f < 6; ++f) s[u][f] = (need >> f) & 1u ? sr[f] : P{};
            o[u] = frames ? fperm[j[u]] : 0;
        }
#pragma unroll
        for (int u = 0; u < PAIR_NT; ++u)
            pair_term<OP, AP, UH, T>(dim, D, A, s[u], o[u], fr, dtf, nr0, nr1, nr2);
    }
    for (; k < e; ++k) {
        const int32_t j = __builtin_nontemporal_load(row + ((int64_t)k << 6));
        const P *sr = rec2(srec, j);
        P s[6];
#pragma unroll
        for (int f = 0; f < 6; ++f) s[f] = (need >> f) & 1u ? sr[f] : P{};
        const int32_t o = frames ? fperm[j] : 0;
        pair_term<OP, AP, UH, T>(dim, D, A, s, o, fr, dtf, nr0, nr1, nr2);
    }
}

template <int OP, bool AP, bool UH, typename T>
__global__ void __launch_bounds__(PAIR_THREADS)
    k_pair(int dim, int64_t nd_total, int64_t nd_real, const int32_t *__restrict__ dperm,
           SrcLists<T> L, Frames F, const T *__restrict__ drec, PairArgs<T> A, int accumulate,
           double *__restrict__ ax, double *__restrict__ ay, double *__restrict__ az,
           double *__restrict__ dt_fac)
{
    constexpr bool DTF = OP == SPH_PAIR_ARTIFICIAL_VISCOSITY || OP == SPH_PAIR_MOMENTUM;
    const int64_t t =
        (int64_t)sph_xcd_tile(blockIdx.x, gridDim.x) * PAIR_THREADS + threadIdx.x;
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
    T dtf = DTF ? T(dt_fac[i]) : T(0.0);
    T o0 = accumulate ? T(ax[i]) : T(0.0);
    T o1 = accumulate ? T(ay[i]) : T(0.0);
    T o2 = (accumulate && az) ? T(az[i]) : T(0.0);
    if (!any) {
        o0 = o0 + T(0.0);
        o1 = o1 + T(0.0);
        o2 = o2 + T(0.0);
    } else {
        const auto *dr = rec2(drec, t);
        const auto d0 = dr[0], d1 = dr[1], d2 = dr[2], d3 = dr[3], d4 = dr[4], d5 = dr[5];
        PairDest<T> D;
        D.xa = d0.x;
        D.ya = d0.y;
        D.za = d1.x;
        D.ma = d1.y;
        D.ua = d2.x;
        D.va = d2.y;
        D.wa = d3.x;
        D.rhoa = d4.x;
        D.ca = d4.y;
        D.ha = d5.x;
        D.Pa = d5.y;
        D.facA = sph_h_dim(D.ha, dim) * sph_cubic_fac1<T>(dim);
        D.rhoa2 = T(1.0) / (D.rhoa * D.rhoa);
        D.wdpA = AP ? sph_cubic_w_fac<T>(A.deltap, T(0.0), T(0.0), D.ha, D.facA) : T(0.0);
        for (int s = 0; s < L.nsrc; ++s) {
            T nr0 = T(0.0), nr1 = T(0.0), nr2 = T(0.0);
            const int32_t *row = L.indices[s] + L.slice_off[s][t >> 6] + (t & 63);
            pair_row<OP, AP, UH, T>(dim, D, A, row, e[s], L.srec[s], F.perm[s], F.f[s], dtf,
                                    nr0, nr1, nr2);
            o0 = o0 + nr0;
            o1 = o1 + nr1;
            o2 = o2 + nr2;
        }
    }
    if (DTF) dt_fac[i] = dtf;
    ax[i] = o0;
    ay[i] = o1;
    if (az) az[i] = o2;
}

template <int OP, bool AP, typename T>
int launch_pair(sph_ctx *ctx, int uniform_h, dim3 grid, int dim, int64_t nd_total,
                int64_t nd_real, const int32_t *dperm, const SrcLists<T> &L, const Frames &F,
                const T *drec, const PairArgs<T> &A, int accumulate, double *ax, double *ay,
                double *az, double *dt_fac)
{
    auto kern = uniform_h ? k_pair<OP, AP, true, T> : k_pair<OP, AP, false, T>;
    hipLaunchKernelGGL(kern, grid, dim3(PAIR_THREADS), 0, ctx->stream, dim, nd_total, nd_real,
                       dperm, L, F, drec, A, accumulate, ax, ay, az, dt_fac);
    SPH_LAUNCH_CHECK(ctx, "k_pair");
    return SPH_OK;
}

template <typename T>
int pair_accel(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real, const int32_t *d_perm,
               const sph_sources *src, const T *d_rec, int uniform_h,
               const sph_pair_params *prm, const sph_wall_frames *frames, int accumulate,
               double *ax, double *ay, double *az, double *dt_fac)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, prm, "sph_pair_accel: null params");
    const int op = prm->op;
    SPH_REQUIRE(ctx, op >= SPH_PAIR_PRESSURE_GRADIENT && op <= SPH_PAIR_BOUNDARY_FORCE,
                "sph_pair_accel: unknown operator %d", op);
    const int ap = prm->with_correction != 0;
    SPH_REQUIRE(ctx, !ap || op == SPH_PAIR_PRESSURE_GRADIENT || op == SPH_PAIR_MOMENTUM,
                "sph_pair_accel: artificial pressure applies to the pressure gradient and "
                "the momentum equation only");
    SPH_REQUIRE(ctx, !ap || prm->deltap >= 0, "sph_pair_accel: deltap cannot be negative");
    SPH_REQUIRE(ctx, dim >= 1 && dim <= 3, "sph_pair_accel: bad dim %d", dim);
    SPH_REQUIRE(ctx, nd_real >= 0 && nd_total >= nd_real, "sph_pair_accel: bad n");
    if (nd_total == 0) return SPH_OK;
    SPH_REQUIRE(ctx, d_rec && ax && ay && d_perm, "sph_pair_accel: null argument");
    const bool dtf = op == SPH_PAIR_ARTIFICIAL_VISCOSITY || op == SPH_PAIR_MOMENTUM;
    SPH_REQUIRE(ctx, !dtf || dt_fac, "sph_pair_accel: this operator needs dt_fac");
    SrcLists<T> L;
    int rc = src_lists<T>(ctx, "sph_pair_accel", src, &L);
    if (rc) return rc;
    Frames F = {};
    if (op == SPH_PAIR_BOUNDARY_FORCE) {
        SPH_REQUIRE(ctx, frames, "sph_pair_accel: the wall force needs tangent/normal frames");
        for (int s = 0; s < src->nsrc; ++s) {
            const double *f[6] = {frames->tx[s], frames->ty[s], frames->tz[s],
                                  frames->nx[s], frames->ny[s], frames->nz[s]};
            SPH_REQUIRE(ctx, frames->s_perm[s], "sph_pair_accel: null frame permutation %d", s);
            F.perm[s] = frames->s_perm[s];
            for (int c = 0; c < 6; ++c) {
                SPH_REQUIRE(ctx, f[c], "sph_pair_accel: null frame component %d of source %d",
                            c, s);
                F.f[s][c] = f[c];
            }
        }
    }
    if (op == SPH_PAIR_PRESSURE_GRADIENT || op == SPH_PAIR_MOMENTUM) {
        rc = sph_prho2_sources(ctx, src, sizeof(T) == sizeof(float));
        if (rc) return rc;
    }
    const PairArgs<T> A = {T(prm->alpha), T(prm->beta), T(prm->eta), T(prm->deltap),
                           T(prm->n),     T(prm->eps),  T(prm->delp)};
    const dim3 grid(sph_blocks(nd_total, PAIR_THREADS));
#define PAIR_ARGS                                                                            \
    ctx, uniform_h, grid, dim, nd_total, nd_real, d_perm, L, F, d_rec, A, accumulate, ax, ay, \
        az, dt_fac
    switch (op) {
    case SPH_PAIR_PRESSURE_GRADIENT:
        return ap ? launch_pair<SPH_PAIR_PRESSURE_GRADIENT, true, T>(PAIR_ARGS)
                  : launch_pair<SPH_PAIR_PRESSURE_GRADIENT, false, T>(PAIR_ARGS);
    case SPH_PAIR_ARTIFICIAL_VISCOSITY:
        return launch_pair<SPH_PAIR_ARTIFICIAL_VISCOSITY, false, T>(PAIR_ARGS);
    case SPH_PAIR_MOMENTUM:
        return ap ? launch_pair<SPH_PAIR_MOMENTUM, true, T>(PAIR_ARGS)
                  : launch_pair<SPH_PAIR_MOMENTUM, false, T>(PAIR_ARGS);
    default:
        return launch_pair<SPH_PAIR_BOUNDARY_FORCE, false, T>(PAIR_ARGS);
    }
#undef PAIR_ARGS
}

}  // namespace

extern "C" int sph_pair_accel(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                              const int32_t *d_perm, const sph_sources *src,
                              const double *d_rec, int uniform_h, const sph_pair_params *prm,
                              const sph_wall_frames *frames, int accumulate, double *ax,
                              double *ay, double *az, double *dt_fac)
{
    return pair_accel<double>(ctx, dim, nd_total, nd_real, d_perm, src, d_rec, uniform_h, prm,
                              frames, accumulate, ax, ay, az, dt_fac);
}

extern "C" int sph_pair_accel_f32(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                                  const int32_t *d_perm, const sph_sources *src,
                                  const float *d_rec, int uniform_h, const sph_pair_params *prm,
                                  const sph_wall_frames *frames, int accumulate, double *ax,
                                  double *ay, double *az, double *dt_fac)
{
    return pair_accel<float>(ctx, dim, nd_total, nd_real, d_perm, src, d_rec, uniform_h, prm,
                             frames, accumulate, ax, ay, az, dt_fac);
}

/*
 * sphref.c -- FP64 CPU oracle for the PySPH particle hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness checker and the
 * CPU baseline: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.  The product path (pysph_amd) never calls it.
 *
 * It restates, in plain C with the same expression order, the Cython
 * arithmetic of the reference (sabago/pysph, PySPH 0.9beta).  Each function
 * cites the reference file:line it follows.  Build flags: -O3
 * -ffp-contract=off (no FMA contraction, like the reference's baseline
 * x86-64/SSE2 build), optional OpenMP for the multi-core baseline.
 *
 * Semantics fixed here (reference citations):
 *   cell id        (int)floor(x / cell_size)          base/cell.pyx:44-84
 *   stencil        (2d+1)^3, d=(int)ceil(r/cell), x outer, y, z inner
 *                                                      base/cell.pyx:95-118, :996-1033
 *   cell contents  ascending particle index           base/cell.pyx:1229-1306
 *   predicate      distance2(src,dst) < (h_dst*rs)^2  base/nnps.pyx:271-356, :525-559
 *                  distance2 order                    base/point.pxd:35-37
 *   self slot      own index swapped to the last slot base/nnps.pyx:596-621
 *   variable h     own row + reversed rows appended   base/nnps.pyx:691-786
 *   cubic spline   W, gradW                           base/kernels.pyx:51-57, :349-453
 *   SPHRho         sum m_b W(r_ab, (h_a+h_b)/2)       sph/funcs/density_funcs.pyx:37-73
 *   Tait           p, cs                              sph/funcs/eos_funcs.pyx:80-129
 *   Momentum + AV  pressure + Monaghan Pi_ab, dt_fac  sph/funcs/pressure_funcs.pyx:185-391
 *   DensityRate    sum (va-vb).gradW m_b               sph/funcs/density_funcs.pyx:116-164
 *   XSPH           sum m_b W/rho_ab (vb-va) eps        sph/funcs/xsph_funcs.pyx:36-91
 *   PressureGrad   -m_b(Pa/ra^2+Pb/rb^2)+art. press.  sph/funcs/pressure_funcs.pyx:19-168
 *   Monaghan AV    Pi_ab gradW, dt_fac (0.01, eta=0)  sph/funcs/viscosity_funcs.pyx:19-157
 *   Momentum deltap artificial pressure branch       sph/funcs/pressure_funcs.pyx:327-365
 *   BoundaryForce  Monaghan wall force (tangent/normal) sph/funcs/boundary_funcs.pyx:9-107
 *   Gravity        += (gx,gy,gz)                      sph/funcs/external_force.pyx:38-44
 *   PositionStep   a_x = u                            sph/funcs/position_funcs.pyx:29-43
 *   stage          x = x0 + a*dt ; PC final 2x - x0   solver/integrator.py:340-367, :632-683
 *   eval loop      out += per-dest result (LocalReal) sph/sph_func.pyx:208-227, :427-441, :513-531
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define REF_PI 3.141592653589793 /* numpy.pi, base/kernels.pyx:23 */

/* ------------------------------------------------------------------ */
/* pow(x, e) of the equation of state and the artificial pressure      */
/* (sph/funcs/eos_funcs.pyx:118-129, pressure_funcs.pyx:121-136)        */
/* ------------------------------------------------------------------ */
/* The reference calls the C library's pow.  The glibc of its time (before
 * 2.28) returned pow correctly rounded; the current one misrounds about one
 * call in a thousand for these arguments.  So for an integral exponent
 * 1 <= e <= 32 and 2^-20 <= |x| <= 2^20 the oracle takes x^e correctly
 * rounded: binary powering in double-double arithmetic (exact products by
 * fma, fast two-sum renormalisation), relative error below 2^-98 before the
 * final rounding.  Other exponents and ranges: the C library's pow.
 * (tests/test_oracle.py checks the rounding against exact rational powers.) */
typedef struct { double hi, lo; } ref_dd;

static ref_dd ref_dd_mul(ref_dd a, ref_dd b)
{
    double p = a.hi * b.hi;
    double e = fma(a.hi, b.hi, -p);
    double t = a.hi * b.lo + a.lo * b.hi;
    double s = e + t;
    double hi = p + s;
    ref_dd r;
    r.hi = hi;
    r.lo = s - (hi - p);
    return r;
}

double sphref_pow(double x, double e)
{
    double ax = fabs(x);
    if (!(e >= 1.0 && e <= 32.0 && e == floor(e) && ax >= 0x1p-20 && ax <= 0x1p20))
        return pow(x, e);
    int n = (int)e;
    ref_dd xd = {x, 0.0}, r = {x, 0.0};
    for (int b = 30 - __builtin_clz((unsigned)n); b >= 0; --b) {
        r = ref_dd_mul(r, r);
        if ((n >> b) & 1) r = ref_dd_mul(r, xd);
    }
    return r.hi;
}

/* ------------------------------------------------------------------ */
/* cells                                                               */
/* ------------------------------------------------------------------ */

/* base/cell.pyx:44-61 */
int sphref_real_to_int(double real_val, double step)
{
    return (int)floor(real_val / step);
}

/* base/cell.pyx:911-941 (radius_scale * max_h, clamped; 0 -> 1.0) */
double sphref_compute_cell_size(double max_h, double radius_scale,
                                double min_size, double max_size)
{
    double cell_size = radius_scale * max_h;
    if (cell_size < min_size)
        cell_size = min_size;
    else if (cell_size > max_size && max_size > 0)
        cell_size = max_size;
    if (cell_size == 0.0)
        cell_size = 1.0;
    return cell_size;
}

typedef struct {
    int cx, cy, cz;
    int32_t idx;
} ref_cell_entry;

static int cmp_entry(const void *a, const void *b)
{
    const ref_cell_entry *p = (const ref_cell_entry *)a;
    const ref_cell_entry *q = (const ref_cell_entry *)b;
    if (p->cx != q->cx) return p->cx < q->cx ? -1 : 1;
    if (p->cy != q->cy) return p->cy < q->cy ? -1 : 1;
    if (p->cz != q->cz) return p->cz < q->cz ? -1 : 1;
    if (p->idx != q->idx) return p->idx < q->idx ? -1 : 1;
    return 0;
}

static int cmp_cell(const ref_cell_entry *p, int cx, int cy, int cz)
{
    if (p->cx != cx) return p->cx < cx ? -1 : 1;
    if (p->cy != cy) return p->cy < cy ? -1 : 1;
    if (p->cz != cz) return p->cz < cz ? -1 : 1;
    return 0;
}

/* first entry with cell >= (cx,cy,cz) */
static int64_t lower_bound(const ref_cell_entry *e, int64_t n, int cx, int cy, int cz)
{
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (cmp_cell(&e[mid], cx, cy, cz) < 0) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* Bin the sources: the dict of cells with ascending indices per cell
 * (base/cell.pyx:1229-1306) restated as entries sorted by (cell, index). */
static ref_cell_entry *bin_sources(const double *sx, const double *sy, const double *sz,
                                   int64_t ns, double cell_size)
{
    ref_cell_entry *e = (ref_cell_entry *)malloc(sizeof(ref_cell_entry) * (size_t)(ns > 0 ? ns : 1));
    for (int64_t j = 0; j < ns; ++j) {
        e[j].cx = sphref_real_to_int(sx[j], cell_size);
        e[j].cy = sphref_real_to_int(sy[j], cell_size);
        e[j].cz = sphref_real_to_int(sz[j], cell_size);
        e[j].idx = (int32_t)j;
    }
    qsort(e, (size_t)ns, sizeof(ref_cell_entry), cmp_entry);
    return e;
}

/* One destination query in reference order: stencil cells x->y->z, each
 * cell's indices ascending, predicate distance2 < radius^2 (strict).
 * base/nnps.pyx:271-356 with radius from :552 and stencil cell.pyx:996-1033.
 * If out != NULL the indices are written; the count is returned. */
static int64_t query_one(const ref_cell_entry *e, int64_t ns,
                         const double *sx, const double *sy, const double *sz,
                         double px, double py, double pz, double radius,
                         double cell_size, int32_t *out)
{
    double radius2 = radius * radius;
    int distance = (int)ceil(radius / cell_size);
    int cx = sphref_real_to_int(px, cell_size);
    int cy = sphref_real_to_int(py, cell_size);
    int cz = sphref_real_to_int(pz, cell_size);
    int64_t cnt = 0;
    for (int i = -distance; i <= distance; ++i)
        for (int j = -distance; j <= distance; ++j)
            for (int k = -distance; k <= distance; ++k) {
                int64_t p = lower_bound(e, ns, cx + i, cy + j, cz + k);
                for (; p < ns && cmp_cell(&e[p], cx + i, cy + j, cz + k) == 0; ++p) {
                    int32_t idx = e[p].idx;
                    /* cPoint_distance2(src, dst), base/point.pxd:35-37 */
                    double d2 = ((sx[idx] - px) * (sx[idx] - px) + (sy[idx] - py) * (sy[idx] - py) +
                                 (sz[idx] - pz) * (sz[idx] - pz));
                    if (d2 < radius2) {
                        if (out) out[cnt] = idx;
                        ++cnt;
                    }
                }
            }
    return cnt;
}

/* Per-destination neighbour lists of FixedDestNbrParticleLocator
 * (base/nnps.pyx:458-621): rows for dests [0, nd_real), indices in the
 * reference order, own index moved to the last slot when same_array.
 * Pass indices == NULL to get the row offsets only (offsets[nd_real+1]);
 * returns nnz.  */
int64_t sphref_nnps(const double *dx, const double *dy, const double *dz, const double *dh,
                    int64_t nd_real,
                    const double *sx, const double *sy, const double *sz, int64_t ns,
                    double radius_scale, double cell_size, int same_array,
                    int64_t *offsets, int32_t *indices, int nthreads)
{
    ref_cell_entry *e = bin_sources(sx, sy, sz, ns, cell_size);
    (void)nthreads;
    if (!indices) {
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads > 0 ? nthreads : 1)
        for (int64_t i = 0; i < nd_real; ++i) {
            double radius = dh[i] * radius_scale;
            offsets[i + 1] = query_one(e, ns, sx, sy, sz, dx[i], dy[i], dz[i], radius, cell_size, NULL);
        }
        offsets[0] = 0;
        for (int64_t i = 0; i < nd_real; ++i) offsets[i + 1] += offsets[i];
    } else {
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads > 0 ? nthreads : 1)
        for (int64_t i = 0; i < nd_real; ++i) {
            double radius = dh[i] * radius_scale;
            int32_t *row = indices + offsets[i];
            int64_t len = query_one(e, ns, sx, sy, sz, dx[i], dy[i], dz[i], radius, cell_size, row);
            /* keep the particle itself at the last index, nnps.pyx:612-619 */
            if (same_array) {
                for (int64_t j = 0; j < len - 1; ++j) {
                    if (row[j] == (int32_t)i) {
                        row[j] = row[len - 1];
                        row[len - 1] = (int32_t)i;
                        break;
                    }
                }
            }
        }
    }
    free(e);
    return offsets[nd_real];
}

/* Symmetric variable-h lists of VarHNbrParticleLocator._update_cache
 * (base/nnps.pyx:742-786):
 *   1. every destination's own query (radius dh[i]*radius_scale, stencil
 *      order), with no move of the own index (the FixedDest swap of
 *      :612-619 is not part of this _update_cache);
 *   2. the reversed FixedDest locator's rows: for every real source j
 *      (j < ns_real, the rows FixedDest fills, :603) the destinations within
 *      sh[j]*radius_scale of source j;
 *   3. for j ascending, for each destination i of j's reversed row, j is
 *      appended to row i unless the row holds it already (LongArray.index).
 * Rows for destinations [0, nd).  Pass indices == NULL for the offsets only;
 * returns nnz. */
typedef struct { int32_t *v; int64_t n, cap; } ref_row;

static void row_push(ref_row *r, int32_t j)
{
    if (r->n == r->cap) {
        r->cap = r->cap ? 2 * r->cap : 16;
        r->v = (int32_t *)realloc(r->v, sizeof(int32_t) * (size_t)r->cap);
    }
    r->v[r->n++] = j;
}

int64_t sphref_nnps_varh(const double *dx, const double *dy, const double *dz, const double *dh,
                         int64_t nd,
                         const double *sx, const double *sy, const double *sz, const double *sh,
                         int64_t ns, int64_t ns_real, double radius_scale, double cell_size,
                         int64_t *offsets, int32_t *indices)
{
    ref_cell_entry *es = bin_sources(sx, sy, sz, ns, cell_size);
    ref_cell_entry *ed = bin_sources(dx, dy, dz, nd, cell_size);
    ref_row *rows = (ref_row *)calloc((size_t)(nd > 0 ? nd : 1), sizeof(ref_row));
    int32_t *buf = (int32_t *)malloc(sizeof(int32_t) * (size_t)((nd > ns ? nd : ns) + 1));
    for (int64_t i = 0; i < nd; ++i) {
        int64_t len = query_one(es, ns, sx, sy, sz, dx[i], dy[i], dz[i], dh[i] * radius_scale,
                                cell_size, buf);
        for (int64_t k = 0; k < len; ++k) row_push(&rows[i], buf[k]);
    }
    for (int64_t j = 0; j < ns_real; ++j) {
        int64_t len = query_one(ed, nd, dx, dy, dz, sx[j], sy[j], sz[j], sh[j] * radius_scale,
                                cell_size, buf);
        for (int64_t k = 0; k < len; ++k) {
            ref_row *r = &rows[buf[k]];
            int found = 0;
            for (int64_t q = 0; q < r->n; ++q)
                if (r->v[q] == (int32_t)j) { found = 1; break; }
            if (!found) row_push(r, (int32_t)j);
        }
    }
    offsets[0] = 0;
    for (int64_t i = 0; i < nd; ++i) {
        offsets[i + 1] = offsets[i] + rows[i].n;
        if (indices)
            for (int64_t q = 0; q < rows[i].n; ++q) indices[offsets[i] + q] = rows[i].v[q];
        free(rows[i].v);
    }
    free(rows);
    free(buf);
    free(es);
    free(ed);
    return offsets[nd];
}

/* Brute-force O(N*M) neighbour sets in ascending index order, the same
 * predicate.  The reference checks its locators against brute force
 * (base/tests/test_locator.py:187-238, base/nnps.pyx:78-94). */
int64_t sphref_nnps_brute(const double *dx, const double *dy, const double *dz, const double *dh,
                          int64_t nd_real,
                          const double *sx, const double *sy, const double *sz, int64_t ns,
                          double radius_scale, int64_t *offsets, int32_t *indices, int nthreads)
{
    (void)nthreads;
    if (!indices) {
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : 1)
        for (int64_t i = 0; i < nd_real; ++i) {
            double radius = dh[i] * radius_scale, r2 = radius * radius;
            int64_t c = 0;
            for (int64_t j = 0; j < ns; ++j) {
                double d2 = ((sx[j] - dx[i]) * (sx[j] - dx[i]) + (sy[j] - dy[i]) * (sy[j] - dy[i]) +
                             (sz[j] - dz[i]) * (sz[j] - dz[i]));
                if (d2 < r2) ++c;
            }
            offsets[i + 1] = c;
        }
        offsets[0] = 0;
        for (int64_t i = 0; i < nd_real; ++i) offsets[i + 1] += offsets[i];
    } else {
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : 1)
        for (int64_t i = 0; i < nd_real; ++i) {
            double radius = dh[i] * radius_scale, r2 = radius * radius;
            int64_t c = offsets[i];
            for (int64_t j = 0; j < ns; ++j) {
                double d2 = ((sx[j] - dx[i]) * (sx[j] - dx[i]) + (sy[j] - dy[i]) * (sy[j] - dy[i]) +
                             (sz[j] - dz[i]) * (sz[j] - dz[i]));
                if (d2 < r2) indices[c++] = (int32_t)j;
            }
        }
    }
    return offsets[nd_real];
}

/* ------------------------------------------------------------------ */
/* cubic spline, base/kernels.pyx:51-57, :349-453                      */
/* ------------------------------------------------------------------ */

static double h_dim(double h, int dim)
{
    if (dim == 1) return 1 / h;
    else if (dim == 2) return 1 / (h * h);
    else return 1 / (h * h * h);
}

/* self.fac = _fac(1.0), base/kernels.pyx:75, :441-447 */
static double cubic_fac1(int dim)
{
    if (dim == 1) return (2. / 3.) / 1.0;
    else if (dim == 2) return 10 / (7 * REF_PI) / (1.0 * 1.0);
    else return 1. / REF_PI / (1.0 * 1.0 * 1.0);
}

/* CubicSplineKernel.function(pa, pb, h), base/kernels.pyx:355-378 */
static double cubic_w(int dim, double pax, double pay, double paz,
                      double pbx, double pby, double pbz, double h)
{
    double rab = sqrt((pax - pbx) * (pax - pbx) + (pay - pby) * (pay - pby) +
                      (paz - pbz) * (paz - pbz));
    double q = rab / h;
    double val;
    double fac = h_dim(h, dim) * cubic_fac1(dim);
    if (q > 2.0)
        val = 0.0;
    else if (q > 1.0)
        val = 0.25 * (2 - q) * (2 - q) * (2 - q);
    else
        val = 1 - 1.5 * (q * q) * (1 - 0.5 * q);
    return val * fac;
}

/* CubicSplineKernel.gradient(pa, pb, h), base/kernels.pyx:380-417 */
static void cubic_gradw(int dim, double pax, double pay, double paz,
                        double pbx, double pby, double pbz, double h, double *g)
{
    double rab = sqrt((pax - pbx) * (pax - pbx) + (pay - pby) * (pay - pby) +
                      (paz - pbz) * (paz - pbz));
    double q = rab / h;
    double val = 0.0;
    double fac = h_dim(h, dim) * cubic_fac1(dim);
    double rx = pax - pbx, ry = pay - pby, rz = paz - pbz;
    if (q > 2.0) {
    } else if (q >= 1.0) {
        val = -0.75 * (2 - q) * (2 - q) / (h * rab);
    } else if (q > 1e-14) {
        val = 3.0 * (0.75 * q - 1) / (h * h);
    }
    g[0] = rx * (val * fac);
    g[1] = ry * (val * fac);
    g[2] = rz * (val * fac);
}

/* vectorised kernel evaluation for tests: W and gradW for n pairs */
void sphref_cubic_spline(int dim, int64_t n,
                         const double *ax, const double *ay, const double *az,
                         const double *bx, const double *by, const double *bz,
                         const double *h, double *w, double *gx, double *gy, double *gz)
{
    for (int64_t i = 0; i < n; ++i) {
        double g[3];
        if (w) w[i] = cubic_w(dim, ax[i], ay[i], az[i], bx[i], by[i], bz[i], h[i]);
        cubic_gradw(dim, ax[i], ay[i], az[i], bx[i], by[i], bz[i], h[i], g);
        if (gx) gx[i] = g[0];
        if (gy) gy[i] = g[1];
        if (gz) gz[i] = g[2];
    }
}

/* ------------------------------------------------------------------ */
/* operators                                                            */
/* ------------------------------------------------------------------ */

/* SPHRho: CSPHFunctionParticle.eval_single + SPHRho.eval_nbr_csph
 * (sph/sph_func.pyx:513-531, sph/funcs/density_funcs.pyx:37-73) under
 * SPHFunction.eval (sph/sph_func.pyx:208-227): out[i] += result. */
void sphref_summation_density(int dim, int64_t nd_real,
                              const double *dx, const double *dy, const double *dz, const double *dh,
                              const double *sx, const double *sy, const double *sz,
                              const double *sh, const double *sm,
                              const int64_t *offsets, const int32_t *indices,
                              double *rho_out, int nthreads)
{
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t i = 0; i < nd_real; ++i) {
        double nr = 0.0, dnr = 0.0;
        for (int64_t k = offsets[i]; k < offsets[i + 1]; ++k) {
            int32_t j = indices[k];
            double ha = dh[i];
            double hb = sh[j];
            double hab = 0.5 * (ha + hb);
            double w = cubic_w(dim, dx[i], dy[i], dz[i], sx[j], sy[j], sz[j], hab);
            nr += w * sm[j];
        }
        if (dnr != 0.0) nr /= dnr;
        rho_out[i] += nr;
    }
}

/* TaitEquation.eval_single, sph/funcs/eos_funcs.pyx:118-129 with
 * B = co*co*ro/gamma (:90); out += result (sph_func.pyx:222-225). */
void sphref_tait(int64_t n_real, const double *rho, double ro, double co, double gamma,
                 double *p_out, double *cs_out)
{
    double B = co * co * ro / gamma;
    for (int64_t i = 0; i < n_real; ++i) {
        double rhoa = rho[i];
        double ratio = rhoa / ro;
        double gamma2 = 0.5 * (gamma - 1.0);
        double tmp = sphref_pow(ratio, gamma);
        p_out[i] += (tmp - 1.0) * B;
        cs_out[i] += sphref_pow(ratio, gamma2) * co;
    }
}

/* MomentumEquation.eval_nbr, sph/funcs/pressure_funcs.pyx:246-391
 * (hks=False, deltap=None) summed by SPHFunctionParticle.eval_single
 * (sph/sph_func.pyx:427-441); out += result; dt_fac max-reduced. */
void sphref_momentum(int dim, int64_t nd_real,
                     const double *dx, const double *dy, const double *dz, const double *dh,
                     const double *drho, const double *dp, const double *du, const double *dv,
                     const double *dw, const double *dcs,
                     const double *sx, const double *sy, const double *sz, const double *sh,
                     const double *sm, const double *srho, const double *sp, const double *su,
                     const double *sv, const double *sw, const double *scs,
                     const int64_t *offsets, const int32_t *indices,
                     double alpha, double beta, double eta,
                     double *ax, double *ay, double *az, double *dt_fac, int nthreads)
{
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t i = 0; i < nd_real; ++i) {
        double nr0 = 0.0, nr1 = 0.0, nr2 = 0.0;
        double dtf = dt_fac[i];
        for (int64_t k = offsets[i]; k < offsets[i + 1]; ++k) {
            int32_t j = indices[k];
            double ha = dh[i];
            double hb = sh[j];
            double hab = 0.5 * (ha + hb);
            double ca = dcs[i];
            double cb = scs[j];
            double rabx = dx[i] - sx[j];
            double raby = dy[i] - sy[j];
            double rabz = dz[i] - sz[j];
            double vabx = du[i] - su[j];
            double vaby = dv[i] - sv[j];
            double vabz = dw[i] - sw[j];
            double dot = vabx * rabx + vaby * raby + vabz * rabz; /* cPoint_dot */
            double rab2 = rabx * rabx + raby * raby + rabz * rabz; /* cPoint_norm */
            double f = fabs(hab * dot / (rab2 + eta * eta * hab * hab));
            if (f > dtf) dtf = f; /* max(d_dt_fac, dt_fac) */
            double Pa = dp[i], rhoa = drho[i];
            double Pb = sp[j], rhob = srho[j], mb = sm[j];
            double rhoa2 = 1.0 / (rhoa * rhoa);
            double rhob2 = 1.0 / (rhob * rhob);
            double tmp = Pa * rhoa2 + Pb * rhob2;
            double piab = 0;
            if (dot < 0) {
                double cab = 0.5 * (ca + cb);
                double rhoab = 0.5 * (rhoa + rhob);
                double mu = hab * dot;
                mu /= (rab2 + eta * eta * hab * hab);
                piab = -alpha * cab * mu + beta * mu * mu;
                piab /= rhoab;
            }
            tmp += piab;
            tmp *= -mb;
            double g[3];
            cubic_gradw(dim, dx[i], dy[i], dz[i], sx[j], sy[j], sz[j], hab, g);
            nr0 += tmp * g[0];
            nr1 += tmp * g[1];
            nr2 += tmp * g[2];
        }
        dt_fac[i] = dtf;
        ax[i] += nr0;
        ay[i] += nr1;
        if (az) az[i] += nr2;
    }
}

/* SPHDensityRate.eval_nbr, sph/funcs/density_funcs.pyx:116-164 (hks off,
 * no kernel corrections): nr += cPoint_dot(va - vb, gradW(dst, src, hab)) * m_b,
 * summed by SPHFunctionParticle.eval_single (sph/sph_func.pyx:427-441);
 * out += result. */
void sphref_density_rate(int dim, int64_t nd_real,
                         const double *dx, const double *dy, const double *dz, const double *dh,
                         const double *du, const double *dv, const double *dw,
                         const double *sx, const double *sy, const double *sz, const double *sh,
                         const double *sm, const double *su, const double *sv, const double *sw,
                         const int64_t *offsets, const int32_t *indices, double *out,
                         int nthreads)
{
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t i = 0; i < nd_real; ++i) {
        double nr = 0.0;
        for (int64_t k = offsets[i]; k < offsets[i + 1]; ++k) {
            int32_t j = indices[k];
            double ha = dh[i];
            double hb = sh[j];
            double hab = 0.5 * (ha + hb);
            double velx = du[i] - su[j];
            double vely = dv[i] - sv[j];
            double velz = dw[i] - sw[j];
            double g[3];
            cubic_gradw(dim, dx[i], dy[i], dz[i], sx[j], sy[j], sz[j], hab, g);
            nr += (velx * g[0] + vely * g[1] + velz * g[2]) * sm[j];
        }
        out[i] += nr;
    }
}

/* XSPHCorrection.eval_nbr_csph, sph/funcs/xsph_funcs.pyx:36-91 (hks off, no
 * Bonnet-Lok correction, so CSPHFunctionParticle.eval_single divides by
 * nothing, sph/sph_func.pyx:513-531): temp = m_b*w/rho_ab,
 * nr += temp*(vb - va)*eps; out += result. */
void sphref_xsph(int dim, int64_t nd_real,
                 const double *dx, const double *dy, const double *dz, const double *dh,
                 const double *drho, const double *du, const double *dv, const double *dw,
                 const double *sx, const double *sy, const double *sz, const double *sh,
                 const double *sm, const double *srho, const double *su, const double *sv,
                 const double *sw, const int64_t *offsets, const int32_t *indices, double eps,
                 double *ox, double *oy, double *oz, int nthreads)
{
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t i = 0; i < nd_real; ++i) {
        double nr0 = 0.0, nr1 = 0.0, nr2 = 0.0;
        for (int64_t k = offsets[i]; k < offsets[i + 1]; ++k) {
            int32_t j = indices[k];
            double ha = dh[i];
            double hb = sh[j];
            double hab = 0.5 * (ha + hb);
            double rhoab = 0.5 * (srho[j] + drho[i]);
            double vbax = su[j] - du[i], vbay = sv[j] - dv[i], vbaz = sw[j] - dw[i];
            double mb = sm[j];
            double w = cubic_w(dim, dx[i], dy[i], dz[i], sx[j], sy[j], sz[j], hab);
            double temp = mb * w / rhoab;
            nr0 += temp * vbax * eps;
            nr1 += temp * vbay * eps;
            nr2 += temp * vbaz * eps;
        }
        ox[i] += nr0;
        if (oy) oy[i] += nr1;
        if (oz) oz[i] += nr2;
    }
}

/* ------------------------------------------------------------------ */
/* split-form operators and wall forces                                 */
/* ------------------------------------------------------------------ */
/* The reference compiles these modules without cdivision, so a float
 * division by zero raises ZeroDivisionError inside a `cdef void eval_nbr`,
 * which Cython reports and ignores: the pair adds nothing from that point on.
 * The divisions that can plausibly meet a zero (W at deltap, the wall
 * distance y, delp, m_a + m_b, h) model that as "skip the rest of the pair". */

/* artificial pressure term of pressure_funcs.pyx:92-140 (hks off): returns 0
 * when the pair raised (W(deltap) == 0), else 1 with *fab set. */
static int art_pressure_fab(int dim, double pax, double pay, double paz, double pbx,
                            double pby, double pbz, double h, double deltap, double n,
                            double *fab)
{
    double w = cubic_w(dim, pax, pay, paz, pbx, pby, pbz, h);
    double wdeltap = cubic_w(dim, deltap, 0.0, 0.0, 0.0, 0.0, 0.0, h);
    if (wdeltap == 0.0) return 0;
    double f = w / wdeltap;
    *fab = sphref_pow(f, n);
    return 1;
}

/* SPHPressureGradient.eval_nbr, sph/funcs/pressure_funcs.pyx:61-151 (hks
 * off, no kernel corrections), summed by SPHFunctionParticle.eval_single
 * (sph/sph_func.pyx:427-441); out += result.  with_correction: deltap given
 * (:27-40). */
void sphref_pressure_gradient(int dim, int64_t nd_real,
                              const double *dx, const double *dy, const double *dz,
                              const double *dh, const double *drho, const double *dp,
                              const double *sx, const double *sy, const double *sz,
                              const double *sh, const double *sm, const double *srho,
                              const double *sp, const int64_t *offsets, const int32_t *indices,
                              int with_correction, double deltap, double n, double eps,
                              double *ax, double *ay, double *az, int nthreads)
{
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t i = 0; i < nd_real; ++i) {
        double nr0 = 0.0, nr1 = 0.0, nr2 = 0.0;
        for (int64_t k = offsets[i]; k < offsets[i + 1]; ++k) {
            int32_t j = indices[k];
            double mb = sm[j], rhoa = drho[i], rhob = srho[j], pa = dp[i], pb = sp[j];
            double ha = dh[i], hb = sh[j];
            double h = 0.5 * (ha + hb);
            double temp = 0.0, artificial_stress = 0.0;
            double rhoa2 = 1.0 / (rhoa * rhoa);
            double rhob2 = 1.0 / (rhob * rhob);
            temp = -mb * (pa * rhoa2 + pb * rhob2);
            if (with_correction) {
                double Ra = pa < 0 ? -eps * pa : 0.0;
                Ra = Ra * rhoa2;
                double Rb = pb < 0 ? -eps * pb : 0.0;
                Rb = Rb * rhob2;
                double fab;
                if (!art_pressure_fab(dim, dx[i], dy[i], dz[i], sx[j], sy[j], sz[j], h, deltap,
                                      n, &fab))
                    continue;
                artificial_stress = mb * (Ra + Rb) * fab;
            }
            temp = temp + artificial_stress;
            double g[3];
            cubic_gradw(dim, dx[i], dy[i], dz[i], sx[j], sy[j], sz[j], h, g);
            nr0 += temp * g[0];
            nr1 += temp * g[1];
            nr2 += temp * g[2];
        }
        ax[i] += nr0;
        ay[i] += nr1;
        if (az) az[i] += nr2;
    }
}

/* MonaghanArtificialViscosity.eval_nbr, sph/funcs/viscosity_funcs.pyx:62-151
 * (hks off).  eta is the function's attribute: its constructor takes an eta
 * argument but never stores it (:30-43), so it is 0 unless set afterwards.
 * dt_fac uses the constant 0.01 (:99).  out += result; dt_fac max-reduced. */
void sphref_artificial_viscosity(int dim, int64_t nd_real,
                                 const double *dx, const double *dy, const double *dz,
                                 const double *dh, const double *drho, const double *du,
                                 const double *dv, const double *dw, const double *dcs,
                                 const double *sx, const double *sy, const double *sz,
                                 const double *sh, const double *sm, const double *srho,
                                 const double *su, const double *sv, const double *sw,
                                 const double *scs, const int64_t *offsets,
                                 const int32_t *indices, double alpha, double beta, double eta,
                                 double *ax, double *ay, double *az, double *dt_fac,
                                 int nthreads)
{
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t i = 0; i < nd_real; ++i) {
        double nr0 = 0.0, nr1 = 0.0, nr2 = 0.0;
        double dtf = dt_fac[i];
        for (int64_t k = offsets[i]; k < offsets[i + 1]; ++k) {
            int32_t j = indices[k];
            double ha = dh[i], hb = sh[j];
            double hab = 0.5 * (ha + hb);
            double rabx = dx[i] - sx[j], raby = dy[i] - sy[j], rabz = dz[i] - sz[j];
            double vabx = du[i] - su[j], vaby = dv[i] - sv[j], vabz = dw[i] - sw[j];
            double ca = dcs[i], cb = scs[j];
            double dot = vabx * rabx + vaby * raby + vabz * rabz;
            double rab2 = rabx * rabx + raby * raby + rabz * rabz;
            double f = fabs(hab * dot / (rab2 + 0.01 * hab * hab));
            if (f > dtf) dtf = f;
            double rhoa = drho[i], rhob = srho[j], mb = sm[j];
            double piab = 0;
            if (dot < 0) {
                double cab = 0.5 * (ca + cb);
                double rhoab = 0.5 * (rhoa + rhob);
                double mu = hab * dot;
                double den = rab2 + eta * eta * hab * hab;
                if (den == 0.0) continue;
                mu /= den;
                piab = -alpha * cab * mu + beta * mu * mu;
                piab /= rhoab;
            }
            double tmp = piab;
            tmp *= -mb;
            double g[3];
            cubic_gradw(dim, dx[i], dy[i], dz[i], sx[j], sy[j], sz[j], hab, g);
            nr0 += tmp * g[0];
            nr1 += tmp * g[1];
            nr2 += tmp * g[2];
        }
        dt_fac[i] = dtf;
        ax[i] += nr0;
        ay[i] += nr1;
        if (az) az[i] += nr2;
    }
}

/* MomentumEquation.eval_nbr with artificial pressure, pressure_funcs.pyx:
 * 246-391 with the deltap branch :327-365 (hks off).  A pair whose W(deltap)
 * is 0 raises after dt_fac was updated: dt_fac keeps the update, the
 * acceleration gets nothing from that pair. */
void sphref_momentum_deltap(int dim, int64_t nd_real,
                            const double *dx, const double *dy, const double *dz,
                            const double *dh, const double *drho, const double *dp,
                            const double *du, const double *dv, const double *dw,
                            const double *dcs, const double *sx, const double *sy,
                            const double *sz, const double *sh, const double *sm,
                            const double *srho, const double *sp, const double *su,
                            const double *sv, const double *sw, const double *scs,
                            const int64_t *offsets, const int32_t *indices, double alpha,
                            double beta, double eta, double deltap, double n, double eps,
                            double *ax, double *ay, double *az, double *dt_fac, int nthreads)
{
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t i = 0; i < nd_real; ++i) {
        double nr0 = 0.0, nr1 = 0.0, nr2 = 0.0;
        double dtf = dt_fac[i];
        for (int64_t k = offsets[i]; k < offsets[i + 1]; ++k) {
            int32_t j = indices[k];
            double ha = dh[i], hb = sh[j];
            double hab = 0.5 * (ha + hb);
            double ca = dcs[i], cb = scs[j];
            double rabx = dx[i] - sx[j], raby = dy[i] - sy[j], rabz = dz[i] - sz[j];
            double vabx = du[i] - su[j], vaby = dv[i] - sv[j], vabz = dw[i] - sw[j];
            double dot = vabx * rabx + vaby * raby + vabz * rabz;
            double rab2 = rabx * rabx + raby * raby + rabz * rabz;
            double f = fabs(hab * dot / (rab2 + eta * eta * hab * hab));
            if (f > dtf) dtf = f;
            double Pa = dp[i], rhoa = drho[i];
            double Pb = sp[j], rhob = srho[j], mb = sm[j];
            double rhoa2 = 1.0 / (rhoa * rhoa);
            double rhob2 = 1.0 / (rhob * rhob);
            double tmp = Pa * rhoa2 + Pb * rhob2;
            double piab = 0;
            if (dot < 0) {
                double cab = 0.5 * (ca + cb);
                double rhoab = 0.5 * (rhoa + rhob);
                double mu = hab * dot;
                mu /= (rab2 + eta * eta * hab * hab);
                piab = -alpha * cab * mu + beta * mu * mu;
                piab /= rhoab;
            }
            tmp += piab;
            tmp *= -mb;
            double Ra = 0.0;
            if (Pa < 0) Ra = -eps * Pa;
            Ra = Ra * rhoa2;
            double Rb = 0.0;
            if (Pb < 0) Rb = -eps * Pb;
            Rb = Rb * rhob2;
            double fab;
            if (!art_pressure_fab(dim, dx[i], dy[i], dz[i], sx[j], sy[j], sz[j], hab, deltap, n,
                                  &fab))
                continue;
            tmp = tmp + mb * (Ra + Rb) * fab;
            double g[3];
            cubic_gradw(dim, dx[i], dy[i], dz[i], sx[j], sy[j], sz[j], hab, g);
            nr0 += tmp * g[0];
            nr1 += tmp * g[1];
            nr2 += tmp * g[2];
        }
        dt_fac[i] = dtf;
        ax[i] += nr0;
        ay[i] += nr1;
        if (az) az[i] += nr2;
    }
}

/* MonaghanBoundaryForce.eval_nbr, sph/funcs/boundary_funcs.pyx:55-107:
 * tangent/normal frames of the wall (source) particle, the destination's h,
 * m and cs. */
void sphref_boundary_force(int64_t nd_real,
                           const double *dx, const double *dy, const double *dz,
                           const double *dh, const double *dm, const double *dcs,
                           const double *sx, const double *sy, const double *sz,
                           const double *sm, const double *stx, const double *sty,
                           const double *stz, const double *snx, const double *sny,
                           const double *snz, const int64_t *offsets, const int32_t *indices,
                           double delp, double *ax, double *ay, double *az, int nthreads)
{
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t i = 0; i < nd_real; ++i) {
        double nr0 = 0.0, nr1 = 0.0, nr2 = 0.0;
        for (int64_t k = offsets[i]; k < offsets[i + 1]; ++k) {
            int32_t j = indices[k];
            double h = dh[i], ma = dm[i], mb = sm[j];
            double nx = snx[j], ny = sny[j], nz = snz[j];
            double tx = stx[j], ty = sty[j], tz = stz[j];
            double cs = dcs[i];
            double rx = dx[i] - sx[j], ry = dy[i] - sy[j], rz = dz[i] - sz[j];
            double x = rx * tx + ry * ty + rz * tz;
            double y = rx * nx + ry * ny + rz * nz;
            double force = 0.0;
            if (h == 0.0) continue;
            double q = y / h;
            if (0 <= fabs(x) && fabs(x) <= delp) {
                if (y == 0.0) continue;
                double beta = 0.02 * cs * cs / y;
                if (delp == 0.0) continue;
                double tforce = 1.0 - fabs(x) / delp;
                double nforce;
                if (0 < q && q <= 2.0 / 3.0)
                    nforce = 2.0 / 3.0;
                else if (2.0 / 3.0 < q && q <= 1.0)
                    nforce = 2 * q * (1.0 - 0.75 * q);
                else if (1.0 < q && q <= 2.0)
                    nforce = 0.5 * (2 - q) * (2 - q);
                else
                    nforce = 0.0;
                if (ma + mb == 0.0) continue;
                force = (mb / (ma + mb)) * nforce * tforce * beta;
            }
            nr0 += force * nx;
            nr1 += force * ny;
            nr2 += force * nz;
        }
        ax[i] += nr0;
        ay[i] += nr1;
        if (az) az[i] += nr2;
    }
}

/* GravityForce.eval_single, sph/funcs/external_force.pyx:38-44 */
void sphref_gravity(int64_t n_real, double gx, double gy, double gz,
                    double *ax, double *ay, double *az)
{
    for (int64_t i = 0; i < n_real; ++i) {
        ax[i] += gx;
        ay[i] += gy;
        if (az) az[i] += gz;
    }
}

/* Integrator.step: updated = initial + step * dt, solver/integrator.py:356-360
 * (numpy: the product is rounded, then the sum). */
void sphref_step(int64_t n, const double *x0, const double *a, double dt, double *x)
{
    for (int64_t i = 0; i < n; ++i) {
        double t = a[i] * dt;
        x[i] = x0[i] + t;
    }
}

/* PredictorCorrectorIntegrator.final_step, solver/integrator.py:632-643 */
void sphref_pc_final(int64_t n, const double *x0, double *x)
{
    for (int64_t i = 0; i < n; ++i) {
        double t = 2 * x[i];
        x[i] = t - x0[i];
    }
}

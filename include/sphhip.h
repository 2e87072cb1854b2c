/*
 * sphhip.h -- C ABI of libsphhip.so, the MI355X (gfx950) particle hot path
 * of PySPH: cell binning + neighbour lists, cubic spline, SummationDensity,
 * Tait EOS, momentum equation with Monaghan artificial viscosity, gravity,
 * position stepping and the integrator stage arithmetic.
 *
 * Conventions
 *  - Every pointer argument is a DEVICE pointer (e.g. torch.Tensor.data_ptr())
 *    unless its comment says "host".  Property arrays are contiguous SoA,
 *    one array per property, in the particle order of the ParticleArray.
 *  - Calls are asynchronous and ordered on the context's stream.  Only the
 *    calls that return a size to the host (sph_grid_build, sph_nnps_build)
 *    synchronise the stream.
 *  - Every entry point returns SPH_OK (0) or a negative status; the message
 *    of the last failure is sph_last_error(ctx).  The library never prints
 *    and never aborts.
 *  - Buffers are owned by the caller; the context owns only scratch (scan
 *    partials, reduction buffers, counters) and frees it in sph_ctx_destroy.
 *  - Accumulating outputs follow SPHFunction.eval (sph/sph_func.pyx:208-227):
 *    with accumulate=0 the output is (0.0 + result), with accumulate=1 it is
 *    (out + result).  Destinations at index >= n_real (non-LocalReal, which
 *    ParticleArray.align_particles keeps at the end) are set to 0.
 */
#ifndef SPHHIP_H
#define SPHHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum sph_status {
    SPH_OK = 0,
    SPH_EINVAL = -1,   /* bad argument */
    SPH_ENEEDCAP = -2, /* output capacity too small; the needed size is returned */
    SPH_EHIP = -3,     /* HIP runtime error */
    SPH_EDOMAIN = -4,  /* cell grid too large / non-finite coordinates */
    SPH_ENOMEM = -5    /* scratch allocation failed */
};

typedef struct sph_ctx sph_ctx;

/* Cell grid shared by every array binned together: the CellManager state
 * that fixes binning (base/cell.pyx:742-1033).  Filled by sph_grid_build. */
typedef struct sph_grid {
    double cell_size;      /* radius_scale*max(h) clamped (cell.pyx:911-941) */
    double max_h, min_h;
    int32_t cmin[3];       /* smallest cell id (int)floor(x/cell_size) per axis */
    int32_t dims[3];       /* cells per axis covering every particle */
    int32_t bits[3];       /* Morton bits per axis, 2^bits[a] >= dims[a] */
    int32_t order;         /* SPH_ORDER_MORTON or SPH_ORDER_LINEAR */
    int64_t ncells;        /* size of the dense cell table (Morton: 2^sum(bits)) */
} sph_grid;

enum { SPH_ORDER_MORTON = 0, SPH_ORDER_LINEAR = 1 };

/* ---- context ---------------------------------------------------------- */

/* Create a context on HIP device `device`; `stream` is a hipStream_t (NULL
 * = the null stream).  Replaces nothing in the reference: the reference's
 * device state lives in pyopencl objects (solver/cl_utils.py:1-110). */
int sph_ctx_create(sph_ctx **out, int device, void *stream);
int sph_ctx_destroy(sph_ctx *ctx);
int sph_ctx_set_stream(sph_ctx *ctx, void *stream);
int sph_ctx_synchronize(sph_ctx *ctx);
const char *sph_last_error(sph_ctx *ctx);
/* library version string (host memory, static) */
const char *sph_version(void);

/* ---- cells and neighbour lists (base/cell.pyx, base/nnps.pyx) ---------- */

/* CellManager.compute_cell_size + bounding box over `narrays` arrays.
 * xs/ys/zs/hs are HOST arrays of DEVICE pointers, ns HOST counts.
 * radius_scale, min_cell_size, max_cell_size as in cell.pyx:802-941.
 * `order` selects the cell order of the counting sort.  Synchronises.
 * Replaces CellManager.compute_cell_size/_compute_minmax_h
 * (base/cell.pyx:911-941, :1140-1158). */
int sph_grid_build(sph_ctx *ctx, int narrays, const double *const *xs, const double *const *ys,
                   const double *const *zs, const double *const *hs, const int64_t *ns,
                   double radius_scale, double min_cell_size, double max_cell_size, int order,
                   int64_t max_cells, sph_grid *grid /* host out */);

/* Counting sort of one array into the grid's cells (stable: ascending
 * particle index inside a cell).  Outputs:
 *   cell_start[grid->ncells + 1]  first sorted position of every cell
 *   perm[n]                       original index at each sorted position
 *   sx, sy, sz [n]                positions in sorted order
 *   cell_of[n]                    cell key of each particle (original order)
 * Replaces CellManager._rebin_particles/_insert_particles
 * (base/cell.pyx:1179-1306). */
int sph_bin_array(sph_ctx *ctx, const sph_grid *grid, const double *x, const double *y,
                  const double *z, int64_t n, int32_t *cell_start, int32_t *perm,
                  double *sx, double *sy, double *sz, int32_t *cell_of);

/* Neighbour lists of FixedDestNbrParticleLocator (base/nnps.pyx:458-621) for
 * destinations [0, nd_real): source j is a neighbour of destination i iff
 * distance2(src_j, dst_i) < (h_i*radius_scale)^2, stencil (2d+1)^3 with
 * d = ceil(h_i*radius_scale/cell_size).  Row i holds source indices (original
 * order) in the reference order: stencil cells x->y->z, ascending index in a
 * cell, and with same_array != 0 the destination's own index moved to the
 * last slot (nnps.pyx:612-619).
 *   offsets[nd_real + 1] (int64), indices[cap] (int32).
 * If cap < nnz, returns SPH_ENEEDCAP with *nnz_out = nnz and offsets filled.
 * Synchronises (reads nnz). */
int sph_nnps_build(sph_ctx *ctx, const sph_grid *grid, const double *dx, const double *dy,
                   const double *dz, const double *dh, int64_t nd_real,
                   const int32_t *src_cell_start, const int32_t *src_perm, const double *src_sx,
                   const double *src_sy, const double *src_sz, int64_t ns, double radius_scale,
                   int same_array, int64_t *offsets, int32_t *indices, int64_t cap,
                   int64_t *nnz_out /* host */);

/* ---- operators --------------------------------------------------------- */

/* SPHRho summation density (sph/funcs/density_funcs.pyx:37-73, driven by
 * CSPHFunctionParticle.eval_single sph/sph_func.pyx:513-531):
 *   rho[i] (+)= sum_{j in row i} m_j W(x_i - x_j, (h_i+h_j)/2),  dim in {1,2,3}.
 * Destinations in [nd_real, nd_total) get 0. */
int sph_summation_density(sph_ctx *ctx, int dim, int64_t nd_real, int64_t nd_total,
                          const int64_t *offsets, const int32_t *indices, const double *dx,
                          const double *dy, const double *dz, const double *dh,
                          const double *sx, const double *sy, const double *sz,
                          const double *sh, const double *sm, int accumulate, double *rho);

/* TaitEquation (sph/funcs/eos_funcs.pyx:80-129):
 *   p (+)= (pow(rho/ro, gamma) - 1) * co*co*ro/gamma,  cs (+)= pow(rho/ro, (gamma-1)/2) * co */
int sph_tait(sph_ctx *ctx, int64_t n_real, int64_t n_total, const double *rho, double ro,
             double co, double gamma, int accumulate, double *p, double *cs);

/* MomentumEquation with Monaghan artificial viscosity, hks=False, no
 * artificial pressure (sph/funcs/pressure_funcs.pyx:185-391):
 *   a_i (+)= sum_j -m_j (P_i/rho_i^2 + P_j/rho_j^2 + Pi_ij) gradW_ij(h_ij)
 *   dt_fac_i = max(dt_fac_i, |h_ij v_ij.r_ij / (r_ij^2 + eta^2 h_ij^2)|)
 * az may be NULL (no z output). dst = {x,y,z,h,rho,p,u,v,w,cs},
 * src = {x,y,z,h,m,rho,p,u,v,w,cs}. */
typedef struct sph_mom_dst {
    const double *x, *y, *z, *h, *rho, *p, *u, *v, *w, *cs;
} sph_mom_dst;
typedef struct sph_mom_src {
    const double *x, *y, *z, *h, *m, *rho, *p, *u, *v, *w, *cs;
} sph_mom_src;
int sph_momentum_av(sph_ctx *ctx, int dim, int64_t nd_real, int64_t nd_total,
                    const int64_t *offsets, const int32_t *indices, const sph_mom_dst *dst,
                    const sph_mom_src *src, double alpha, double beta, double eta,
                    int accumulate, double *ax, double *ay, double *az, double *dt_fac);

/* GravityForce (sph/funcs/external_force.pyx:38-44): a (+)= g; az may be NULL. */
int sph_gravity(sph_ctx *ctx, int64_t n_real, int64_t n_total, double gx, double gy, double gz,
                int accumulate, double *ax, double *ay, double *az);

/* PositionStepping (sph/funcs/position_funcs.pyx:29-43): a_x = u (assignment). */
int sph_position_stepping(sph_ctx *ctx, int64_t n_real, int64_t n_total, const double *u,
                          const double *v, const double *w, double *ax, double *ay, double *az);

/* ---- integrator stage arithmetic (solver/integrator.py) ----------------- */

#define SPH_MAX_STEP_PROPS 16
typedef struct sph_step_props {
    int32_t nprops;
    double *x[SPH_MAX_STEP_PROPS];         /* stepped property (out) */
    const double *x0[SPH_MAX_STEP_PROPS];  /* initial value */
    const double *a[SPH_MAX_STEP_PROPS];   /* acceleration */
} sph_step_props;

/* Integrator.step (integrator.py:340-367): x = x0 + a*dt for every prop. */
int sph_stage_step(sph_ctx *ctx, int64_t n, const sph_step_props *props, double dt);
/* PredictorCorrectorIntegrator.final_step (integrator.py:632-643): x = 2x - x0 */
int sph_pc_final(sph_ctx *ctx, int64_t n, const sph_step_props *props);
/* copy n doubles (save_initial_arrays, integrator.py:282-285) */
int sph_copy(sph_ctx *ctx, int64_t n, const double *src, double *dst);
/* fill n doubles with value (reset_accelerations / set_to_zero) */
int sph_fill(sph_ctx *ctx, int64_t n, double value, double *dst);

#ifdef __cplusplus
}
#endif

#endif /* SPHHIP_H */

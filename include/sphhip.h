/*
 * sphhip.h -- C ABI of libsphhip.so, the MI355X (gfx950) particle hot path
 * of PySPH: cell binning + neighbour lists, cubic spline, SummationDensity,
 * Tait EOS, momentum equation with Monaghan artificial viscosity, gravity,
 * position stepping and the integrator stage arithmetic.
 *
 * Conventions
 *  - Every pointer argument is a DEVICE pointer (e.g. torch.Tensor.data_ptr())
 *    unless its comment says "host".  Property arrays are contiguous SoA,
 *    one array per property.
 *  - Two particle orders exist.  ORIGINAL order is the order of the caller's
 *    ParticleArray.  SORTED order is the order of sph_bin_array: particles
 *    grouped by cell (cells in Morton or linear order), ascending original
 *    index inside a cell; perm[t] is the original index at sorted position
 *    t and rank[i] the sorted position of original index i.  "Shadow"
 *    arrays are properties copied into sorted order (sph_gather_sorted);
 *    the neighbour-list and operator kernels read shadows (coalesced,
 *    cell-coherent loads) and write their results in original order (and
 *    optionally into a shadow too).
 *  - Calls are asynchronous and ordered on the context's stream.  Only the
 *    calls that return a size to the host (sph_grid_build, sph_nnps_build)
 *    synchronise the stream.
 *  - Every entry point returns SPH_OK (0) or a negative status; the message
 *    of the last failure is sph_last_error(ctx).  The library never prints
 *    and never aborts.
 *  - Buffers are owned by the caller; the context owns only scratch (scan
 *    partials, reduction buffers, key tables) and frees it in sph_ctx_destroy.
 *  - Accumulating outputs follow SPHFunction.eval (sph/sph_func.pyx:208-227):
 *    with accumulate=0 the output is (0.0 + result), with accumulate=1 it is
 *    (out + result).  Destinations whose original index is >= n_real
 *    (non-LocalReal, which ParticleArray.align_particles keeps at the end)
 *    get 0.
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

#define SPH_MAX_PROPS 16
#define SPH_MAX_SRC 4 /* source arrays of one multi-source operator call */

/* ---- context ---------------------------------------------------------- */

/* Create a context on HIP device `device`; `stream` is a hipStream_t (NULL
 * = the null stream).  The reference keeps its device state in pyopencl
 * objects (solver/cl_utils.py:1-110, particle_array.pyx:1269-1396). */
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
 * `order` selects the cell order of the counting sort; max_cells bounds the
 * dense cell table; `pad` empty cells are added on every side of the
 * bounding box (cell ids floor(x/cell_size) do not depend on it, so neither
 * do the lists).  Synchronises.  Replaces
 * CellManager.compute_cell_size/_compute_minmax_h (cell.pyx:911-941, :1140-1158). */
int sph_grid_build(sph_ctx *ctx, int narrays, const double *const *xs, const double *const *ys,
                   const double *const *zs, const double *const *hs, const int64_t *ns,
                   double radius_scale, double min_cell_size, double max_cell_size, int order,
                   int64_t max_cells, int pad, sph_grid *grid /* host out */);

/* Whether a grid can be kept for the next binning, without synchronising:
 * *flags (device int32) = OR over the particles of the arrays of 1 (within
 * `margin` cells of the grid's edge: rebuild the grid before the next
 * binning), 2 (outside the grid: a binning with it is invalid) and 8 (a
 * non-finite coordinate).  The cell-manager counterpart of the reference's
 * re-binning after a move (cell.pyx:943-976), which recomputes the grid
 * every time; the caller keeps the padded grid while h is unchanged and the
 * flags stay clear. */
int sph_grid_check(sph_ctx *ctx, int narrays, const double *const *xs, const double *const *ys,
                   const double *const *zs, const int64_t *ns, const sph_grid *grid, int margin,
                   int32_t *flags);

/* Counting sort of one array into the grid's cells (stable: ascending
 * original index inside a cell).  Outputs:
 *   cell_start[grid->ncells + 1]  first sorted position of every cell
 *   perm[n]                       original index at each sorted position
 *   rank[n]                       sorted position of each original index
 *   cell_of[n]                    cell key of each particle (original order)
 *   cells[min(n, ncells)]         keys of the non-empty cells, ascending
 *   *ncells_out (host)            number of non-empty cells
 * Synchronises (reads the cell count).  Replaces
 * CellManager._rebin_particles/_insert_particles (base/cell.pyx:1179-1306). */
int sph_bin_array(sph_ctx *ctx, const sph_grid *grid, const double *x, const double *y,
                  const double *z, int64_t n, int32_t *cell_start, int32_t *perm, int32_t *rank,
                  int32_t *cell_of, int32_t *cells, int64_t *ncells_out);
/* The same without a host synchronisation: the number of non-empty cells is
 * written to the device word *ncells_dev (for sph_nnps_build_dev). */
int sph_bin_array_dev(sph_ctx *ctx, const sph_grid *grid, const double *x, const double *y,
                      const double *z, int64_t n, int32_t *cell_start, int32_t *perm,
                      int32_t *rank, int32_t *cell_of, int32_t *cells, int32_t *ncells_dev);

/* Sorted particle records: SPH_REC doubles per particle in sorted order,
 * the layout every neighbour kernel reads (one 96-byte record per
 * neighbour instead of one load per property).  Fields are paired in
 * 16-byte slots by use: (x y)(z m) is all SPHRho reads when h is uniform,
 * (u v)(w P/rho^2)(rho cs) complete MomentumEquation, (h p) last. */
#define SPH_REC 12
enum sph_rec_field {
    SPH_R_X = 0, SPH_R_Y, SPH_R_Z, SPH_R_M, SPH_R_U, SPH_R_V, SPH_R_W,
    SPH_R_PRHO2, /* p * (1/(rho*rho)), filled by sph_momentum_av or sph_tait */
    SPH_R_RHO, SPH_R_CS, SPH_R_H, SPH_R_P
};

/* Shadows: dst[k][t * dst_stride] = src[k][perm[t]] for k < nprops (sorted
 * copies of ParticleArray properties, e.g. record fields with stride
 * SPH_REC).  No reference counterpart: the reference reads its carrays in
 * place (sph_func.pyx:180-206). */
typedef struct sph_prop_list {
    int32_t nprops;
    int64_t dst_stride;
    const double *src[SPH_MAX_PROPS];
    double *dst[SPH_MAX_PROPS];
} sph_prop_list;
int sph_gather_sorted(sph_ctx *ctx, int64_t n, const int32_t *perm, const sph_prop_list *props);

/* ---- single precision (FP32) variant ----------------------------------
 * The reference runs its OpenCL path in single precision when the arrays
 * are created with cl_precision="single" (base/particle_array.pyx:1400-1488,
 * sph/sph_calc.pyx:444-714).  Here the particle arrays stay FP64; the
 * sorted records are float and the list predicate and operator arithmetic
 * run in float.  Pairs at d ~ 2h may differ from the FP64 lists. */
typedef struct sph_prop_list_f32 {
    int32_t nprops;
    int64_t dst_stride;
    const double *src[SPH_MAX_PROPS];
    float *dst[SPH_MAX_PROPS];
} sph_prop_list_f32;
/* dst[k][t * dst_stride] = (float)src[k][perm[t]] */
int sph_gather_sorted_f32(sph_ctx *ctx, int64_t n, const int32_t *perm,
                          const sph_prop_list_f32 *props);
/* The same records written from the original order with rank (original ->
 * sorted): dst[k][rank[i] * dst_stride] = src[k][i] (coalesced reads; the
 * faster direction when many fields are refreshed).  Same results as the
 * gathers. */
int sph_scatter_sorted(sph_ctx *ctx, int64_t n, const int32_t *rank, const sph_prop_list *props);
/* Two sph_scatter_sorted calls over one rank in one pass (each list with its
 * own dst_stride: the records and another sorted layout of some of the same
 * fields); same results as the two calls.  No counterpart in the reference,
 * which keeps no sorted copies (base/cell.pyx:1179-1306 bins indices only). */
int sph_scatter_sorted2(sph_ctx *ctx, int64_t n, const int32_t *rank, const sph_prop_list *props,
                        const sph_prop_list *props2);
int sph_scatter_sorted_f32(sph_ctx *ctx, int64_t n, const int32_t *rank,
                           const sph_prop_list_f32 *props);

/* Neighbour lists of FixedDestNbrParticleLocator (base/nnps.pyx:458-621)
 * in SORTED space, stored as sliced ELLPACK with 64-row slices (SELL-64):
 * destinations are grouped into slices of 64 consecutive sorted positions;
 * entry k of the row of sorted destination t is
 *     indices[slice_off[t / 64] + 64 * k + t % 64],   k < counts[t],
 * so the 64 lanes of a wavefront read one coalesced 256-byte line per k.
 * Rows are empty when d_perm[t] >= nd_real; entries are SORTED source
 * positions.  Source j is a neighbour of destination i iff
 * distance2(src_j, dst_i) < (h_i*radius_scale)^2 (stencil (2d+1)^3,
 * d = ceil(h_i*radius_scale/cell_size)).  Entries follow the reference
 * order: stencil cells x -> y -> z, ascending index inside a cell, and with
 * same_array != 0 the destination itself moved to the last slot
 * (nnps.pyx:612-619).  One wavefront handles one destination cell; its
 * lanes test 64 candidates at a time and compact the hits with ballots.
 *   d_rec [nd_total*SPH_REC], d_perm [nd_total], d_cell_start and d_cells /
 *   nd_cells: the destination's binning; s_cell_start, s_rec [ns*SPH_REC]:
 *   the source's.
 *   counts[nd_total] (int32), slice_off[ceil(nd_total/64) + 1] (int64),
 *   indices[cap] (int32).
 * Returns *slots_out = slice_off[last] (the storage the entries need) and
 * *nnz_out = sum of counts; if cap < slots returns SPH_ENEEDCAP with both
 * filled.  Synchronises (reads the sizes). */
int sph_nnps_build(sph_ctx *ctx, const sph_grid *grid, const double *d_rec,
                   const int32_t *d_perm, const int32_t *d_cell_start, const int32_t *d_cells,
                   int64_t nd_cells, int64_t nd_total, int64_t nd_real,
                   const int32_t *s_cell_start, const double *s_rec, int64_t ns,
                   double radius_scale, int same_array, int32_t *counts, int64_t *slice_off,
                   int32_t *indices, int64_t cap, int64_t *slots_out /* host */,
                   int64_t *nnz_out /* host */);

/* The same build without a host synchronisation: the destination's count of
 * non-empty cells is read on the device (nd_cells_dev, from
 * sph_bin_array_dev; max_cells bounds it and sizes the launch), the totals
 * go to totals[3] = {slots, nnz, status} (device) and the capacity checks to the
 * device word *status: bit 0 when cap < slots (no entry written, every row
 * empty), bit 4 when more than max_cells cells are non-empty (the cells past
 * the bound get no rows).  totals[2] receives the status word after the
 * build, so one copy of totals[3] = {slots, nnz, status} tells the caller
 * everything; it must discard the evaluation that used a build with a
 * status bit set. */
int sph_nnps_build_dev(sph_ctx *ctx, const sph_grid *grid, const double *d_rec,
                       const int32_t *d_perm, const int32_t *d_cell_start, const int32_t *d_cells,
                       const int32_t *nd_cells_dev, int64_t max_cells, int64_t nd_total,
                       int64_t nd_real, const int32_t *s_cell_start, const double *s_rec,
                       int64_t ns, double radius_scale, int same_array, int32_t *counts,
                       int64_t *slice_off, int32_t *indices, int64_t cap, int64_t *totals,
                       int32_t *status, const uint8_t *s_near, int near_reach);

/* Flags of the cells within `reach` cells (Chebyshev) of a non-empty cell of
 * one array: near[key] = 1 for those, 0 for every other key of the grid
 * (grid->ncells bytes, device).  The array's binning: s_cell_start, the
 * list s_cells of its non-empty cells with their count in the device word
 * *ns_cells_dev (bounded by ns_cells; NULL: ns_cells is the count), its
 * permutation s_perm and its double coordinates in original order.  Passed
 * as s_near with near_reach = reach to the device builds (cross-array lists
 * only), it lets the count pass skip a destination cell with no source cell
 * near after two loads: a table built once serves every build until that
 * array is re-binned (walls).  The caller guarantees that no destination's
 * stencil reach (int)ceil(h*radius_scale/cell_size) exceeds near_reach
 * (e.g. from grid->max_h when no h changed since the grid was built); the
 * cells are skipped on the flag alone.  With s_near NULL the builds read
 * each stencil's cell ranges instead.  No counterpart in the reference,
 * whose cell lists hold one array's cells (base/cell.pyx:742-1643). */
int sph_cell_near(sph_ctx *ctx, const sph_grid *grid, const int32_t *s_cell_start,
                  const int32_t *s_cells, const int32_t *ns_cells_dev, int64_t ns_cells,
                  const int32_t *s_perm, const double *s_x, const double *s_y,
                  const double *s_z, int reach, uint8_t *near);

/* sph_nnps_build over float records (d_rec, s_rec with SPH_REC floats per
 * particle): the distance test d2 < (h*radius_scale)^2 is evaluated in
 * float.  d_x, d_y, d_z are the destination's double coordinates in
 * original order (the ones the binning used), from which each destination
 * cell's position is taken. */
int sph_nnps_build_f32(sph_ctx *ctx, const sph_grid *grid, const float *d_rec,
                       const double *d_x, const double *d_y, const double *d_z,
                       const int32_t *d_perm, const int32_t *d_cell_start,
                       const int32_t *d_cells, int64_t nd_cells, int64_t nd_total,
                       int64_t nd_real, const int32_t *s_cell_start, const float *s_rec,
                       int64_t ns, double radius_scale, int same_array, int32_t *counts,
                       int64_t *slice_off, int32_t *indices, int64_t cap, int64_t *slots_out,
                       int64_t *nnz_out);
int sph_nnps_build_f32_dev(sph_ctx *ctx, const sph_grid *grid, const float *d_rec,
                           const double *d_x, const double *d_y, const double *d_z,
                           const int32_t *d_perm, const int32_t *d_cell_start,
                           const int32_t *d_cells, const int32_t *nd_cells_dev,
                           int64_t max_cells, int64_t nd_total, int64_t nd_real,
                           const int32_t *s_cell_start, const float *s_rec, int64_t ns,
                           double radius_scale, int same_array, int32_t *counts,
                           int64_t *slice_off, int32_t *indices, int64_t cap, int64_t *totals,
                           int32_t *status, const uint8_t *s_near, int near_reach);

/* The arguments of one sph_nnps_build_dev (f32 = 0: d_rec and s_rec are
 * double records, d_x/d_y/d_z unused) or sph_nnps_build_f32_dev (f32 = 1),
 * with the meaning they have there. */
typedef struct sph_nnps_dev_args {
    int32_t f32;
    int32_t same_array;
    const void *d_rec;
    const double *d_x, *d_y, *d_z;
    const int32_t *d_perm, *d_cell_start, *d_cells, *nd_cells_dev;
    int64_t max_cells, nd_total, nd_real;
    const int32_t *s_cell_start;
    const void *s_rec;
    int64_t ns;
    double radius_scale;
    int32_t *counts;
    int64_t *slice_off;
    int32_t *indices;
    int64_t cap;
    int64_t *totals;
    int32_t *status;
    const uint8_t *s_near;
    int32_t near_reach;
} sph_nnps_dev_args;

/* Two device-side builds over one grid (the lists of one destination from
 * two source arrays, say) with the results of two sph_nnps_build(_f32)_dev
 * calls, their count passes and their fills each in one launch: the smaller
 * build's blocks run in the tail of the larger one's.  Both builds have one
 * precision and no output buffer in common.  The reference builds each
 * locator's cache on its own (base/nnps.pyx:566-660). */
int sph_nnps_build_dev_pair(sph_ctx *ctx, const sph_grid *grid, const sph_nnps_dev_args *b0,
                            const sph_nnps_dev_args *b1);

/* Symmetric variable-h lists of VarHNbrParticleLocator (base/nnps.pyx:
 * 691-786, chosen by NNPSManager(variable_h=True), :895-910), in the
 * SELL-64 layout of sph_nnps_build.  Row of sorted destination t (LocalReal
 * only; other rows empty): the sources within h_t*radius_scale in the
 * stencil order of sph_nnps_build with the own index NOT moved last, then
 * the real sources j (original index < ns_real) outside that radius whose
 * own radius h_j*radius_scale reaches t, in ascending original index -- the
 * order in which the reference appends the reversed locator's rows
 * (:742-786).  FP64 records with h (d_rec, s_rec); s_perm / s_rank: the
 * source's sorted->original and original->sorted permutations.  max_h: a
 * bound of every h of both arrays (stencil reach ceil(max_h*rs/cell) <= 256).
 * The destination's binning as for sph_nnps_build: d_cell_start, the list
 * d_cells of its non-empty cells, and their count nd_cells, or, with
 * nd_cells_dev non-NULL, the count on the device and nd_cells a bound of it
 * (a larger count is an error).  One wavefront per destination cell tests
 * the cell's candidates (the stencil box of the largest reach) against its
 * destinations; cells whose box exceeds 64 cells or 1024 candidates get
 * one lane per destination.  Reads the totals back (synchronises); returns
 * SPH_ENEEDCAP with *slots_out filled when cap is too small (nothing
 * written). */
int sph_nnps_build_varh(sph_ctx *ctx, const sph_grid *grid, const double *d_rec,
                        const int32_t *d_perm, const int32_t *d_cell_start,
                        const int32_t *d_cells, const int32_t *nd_cells_dev, int64_t nd_cells,
                        int64_t nd_total, int64_t nd_real, const int32_t *s_cell_start,
                        const double *s_rec, const int32_t *s_perm, const int32_t *s_rank,
                        int64_t ns, int64_t ns_real, double radius_scale, double max_h,
                        int32_t *counts, int64_t *slice_off, int32_t *indices, int64_t cap,
                        int64_t *slots_out, int64_t *nnz_out);
/* The same lists over float records (the single-precision variant,
 * cl_precision="single"): both distance tests in float; d_x, d_y, d_z are
 * the destination's double coordinates in original order (the ones the
 * binning used), from which each destination's cell is taken.  Pairs at a
 * distance within float rounding of either radius may differ from the FP64
 * lists. */
int sph_nnps_build_varh_f32(sph_ctx *ctx, const sph_grid *grid, const float *d_rec,
                            const double *d_x, const double *d_y, const double *d_z,
                            const int32_t *d_perm, const int32_t *d_cell_start,
                            const int32_t *d_cells, const int32_t *nd_cells_dev,
                            int64_t nd_cells, int64_t nd_total, int64_t nd_real,
                            const int32_t *s_cell_start, const float *s_rec,
                            const int32_t *s_perm, const int32_t *s_rank, int64_t ns,
                            int64_t ns_real, double radius_scale, double max_h, int32_t *counts,
                            int64_t *slice_off, int32_t *indices, int64_t cap,
                            int64_t *slots_out, int64_t *nnz_out);

/* The same lists in ORIGINAL index space as CSR, the layout of the
 * reference's per-destination LongArrays (nnps.pyx:596-621): row i
 * (i < nd_real) holds original source indices in the same order.
 * out_offsets[nd_real + 1]; out_indices[nnz]. */
int sph_nnps_export(sph_ctx *ctx, int64_t nd_real, const int32_t *d_rank,
                    const int32_t *s_perm, const int32_t *counts, const int64_t *slice_off,
                    const int32_t *indices, int64_t *out_offsets, int32_t *out_indices);

/* ---- operators (sorted records in, original order out) ------------------ */

/* SPHRho summation density (sph/funcs/density_funcs.pyx:37-73, driven by
 * CSPHFunctionParticle.eval_single sph/sph_func.pyx:513-531):
 *   rho_i (+)= sum_{j in row} m_j W(x_i - x_j, (h_i+h_j)/2),  dim in {1,2,3}.
 * (counts, slice_off, indices) are the SELL-64 lists of sph_nnps_build.
 * rho[nd_total] is written in original order (rho[d_perm[t]]); out_s (may
 * be NULL) receives the same values at out_s[t * out_s_stride].
 * uniform_h = 1 asserts that every destination and source particle has the
 * same h (then (h_i+h_j)/2 == h_i exactly and h_j is not read). */
int sph_summation_density(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                          const int32_t *d_perm, const int32_t *counts,
                          const int64_t *slice_off, const int32_t *indices,
                          const double *d_rec, const double *s_rec, int uniform_h,
                          int accumulate, double *rho, double *out_s, int64_t out_s_stride);

/* TaitEquation (sph/funcs/eos_funcs.pyx:80-129) over sorted positions:
 *   p (+)= (pow(rho/ro, gamma) - 1) * co*co*ro/gamma,
 *   cs (+)= pow(rho/ro, (gamma-1)/2) * co
 * rho_s[t * s_stride] is the sorted density; p, cs are written in original
 * order and p_s, cs_s (may be NULL) at [t * s_stride].  prho2_s (may be
 * NULL) receives P/rho^2 = p * (1/(rho*rho)) of the final p at the same
 * stride, the per-source term of MomentumEquation (pressure_funcs.pyx:302-305),
 * so that sph_momentum_av can skip its own pass over the sources. */
int sph_tait(sph_ctx *ctx, int64_t n_total, int64_t n_real, const int32_t *perm,
             const double *rho_s, int64_t s_stride, double ro, double co, double gamma,
             int accumulate, double *p, double *cs, double *p_s, double *cs_s,
             double *prho2_s);

/* MomentumEquation with Monaghan artificial viscosity, hks=False, no
 * artificial pressure (sph/funcs/pressure_funcs.pyx:185-391):
 *   a_i (+)= sum_j -m_j (P_i/rho_i^2 + P_j/rho_j^2 + Pi_ij) gradW_ij(h_ij)
 *   dt_fac_i = max(dt_fac_i, |h_ij v_ij.r_ij / (r_ij^2 + eta^2 h_ij^2)|)
 * d_rec/s_rec are sorted records (s_rec has ns records; its SPH_R_PRHO2
 * field is recomputed from its p and rho fields first -- ns = 0 when the
 * caller knows it is current, e.g. written by sph_tait); ax, ay, az (az may
 * be NULL) and dt_fac are in original order.  uniform_h as for
 * sph_summation_density. */
int sph_momentum_av(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                    const int32_t *d_perm, const int32_t *counts, const int64_t *slice_off,
                    const int32_t *indices,
                    const double *d_rec, double *s_rec, int64_t ns, int uniform_h, double alpha,
                    double beta, double eta, int accumulate, double *ax, double *ay, double *az,
                    double *dt_fac);

/* Single-precision twins of the three operators: float records, float
 * arithmetic, results widened to the double original-order arrays (out_s,
 * p_s, cs_s are float record fields). */
int sph_summation_density_f32(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                              const int32_t *d_perm, const int32_t *counts,
                              const int64_t *slice_off, const int32_t *indices,
                              const float *d_rec, const float *s_rec, int uniform_h,
                              int accumulate, double *rho, float *out_s, int64_t out_s_stride);
int sph_tait_f32(sph_ctx *ctx, int64_t n_total, int64_t n_real, const int32_t *perm,
                 const float *rho_s, int64_t s_stride, double ro, double co, double gamma,
                 int accumulate, double *p, double *cs, float *p_s, float *cs_s,
                 float *prho2_s);
int sph_momentum_av_f32(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                        const int32_t *d_perm, const int32_t *counts, const int64_t *slice_off,
                        const int32_t *indices, const float *d_rec, float *s_rec, int64_t ns,
                        int uniform_h, double alpha, double beta, double eta, int accumulate,
                        double *ax, double *ay, double *az, double *dt_fac);

/* Every function of one calc (SPHCalc.sph, sph/sph_calc.pyx:307-358: one
 * function per source array, e.g. fluid and walls) in one launch: per
 * destination each source's row is summed from 0 in its list order and added
 * to the output in source order, out = ((base + r_1) + r_2) ..., the
 * arithmetic of consecutive single-source calls with accumulate set from the
 * second on.  counts/slice_off/indices: each source's SELL-64 lists for this
 * destination; s_rec: its sorted records (double*, or float* for _f32);
 * ns: momentum only, records whose P/rho^2 field is refreshed first (0 when
 * it is current); s_pm: optional compact copy of the records' first four
 * fields (x, y, z, m; 4 values per sorted particle, the records' type), read
 * by the density gathers instead of those fields (fewer cache lines per
 * gather); NULL, and every other operator, reads them from s_rec. */
typedef struct sph_sources {
    int32_t nsrc;
    const int32_t *counts[SPH_MAX_SRC];
    const int64_t *slice_off[SPH_MAX_SRC];
    const int32_t *indices[SPH_MAX_SRC];
    const void *s_rec[SPH_MAX_SRC];
    int64_t ns[SPH_MAX_SRC];
    const void *s_pm[SPH_MAX_SRC];
    /* the _staged operators only: each source's cell starts (its binning's
     * cell_start) and whether it is the destination array itself */
    const int32_t *s_cell_start[SPH_MAX_SRC];
    int32_t same[SPH_MAX_SRC];
} sph_sources;
int sph_summation_density_multi(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                                const int32_t *d_perm, const sph_sources *src,
                                const double *d_rec, int uniform_h, int accumulate,
                                double *rho, double *out_s, int64_t out_s_stride);
int sph_summation_density_multi_f32(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                                    const int32_t *d_perm, const sph_sources *src,
                                    const float *d_rec, int uniform_h, int accumulate,
                                    double *rho, float *out_s, int64_t out_s_stride);
/* sph_summation_density_multi followed by sph_tait on the same destination
 * array, in one launch: the density calc and the equation-of-state calc of
 * one evaluation (SPHCalc.sph of SPHRho then of TaitEquation,
 * sph/sph_calc.pyx:405-440, funcs/density_funcs.pyx:37-73,
 * funcs/eos_funcs.pyx:118-129).  Results equal the two calls in turn bit for
 * bit: Tait takes the density just written to the record (out_s, required),
 * and its record fields p_s, cs_s, prho2_s share the record stride. */
int sph_summation_density_tait(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                               const int32_t *d_perm, const sph_sources *src,
                               const double *d_rec, int uniform_h, int accumulate,
                               double *rho, double *out_s, int64_t out_s_stride, double ro,
                               double co, double gamma, int eos_accumulate, double *p,
                               double *cs, double *p_s, double *cs_s, double *prho2_s);
int sph_summation_density_tait_f32(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                                   const int32_t *d_perm, const sph_sources *src,
                                   const float *d_rec, int uniform_h, int accumulate,
                                   double *rho, float *out_s, int64_t out_s_stride, double ro,
                                   double co, double gamma, int eos_accumulate, double *p,
                                   double *cs, float *p_s, float *cs_s, float *prho2_s);
int sph_momentum_av_multi(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                          const int32_t *d_perm, const sph_sources *src, const double *d_rec,
                          int uniform_h, double alpha, double beta, double eta, int accumulate,
                          double *ax, double *ay, double *az, double *dt_fac);
int sph_momentum_av_multi_f32(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                              const int32_t *d_perm, const sph_sources *src, const float *d_rec,
                              int uniform_h, double alpha, double beta, double eta,
                              int accumulate, double *ax, double *ay, double *az,
                              double *dt_fac);

/* The same two operators (FP64) with the neighbour records staged in LDS,
 * for grids in the linear cell order (grid->order == SPH_ORDER_LINEAR).
 * With linear keys the reference's row order (stencil x -> y -> z, ascending
 * index, own index last) is ascending sorted position apart from the
 * own-index swap, and each stencil pencil (cx+dx, cy+dy, all z of the reach)
 * is one contiguous range of sorted positions: one wavefront per 64-row slice
 * copies the pencil's records into LDS with contiguous loads and sums its
 * rows' entries from there, pencil by pencil, in list order.  Results equal
 * sph_summation_density_multi / sph_momentum_av_multi bit for bit (the same
 * sums in the same order).  Extra arguments: the grid the lists were built
 * on, their radius_scale, and src->s_cell_start / src->same.  Same loop as
 * SPHFunction.eval over each source (sph/sph_func.pyx:208-227). */
int sph_summation_density_staged(sph_ctx *ctx, const sph_grid *grid, double radius_scale,
                                 int dim, int64_t nd_total, int64_t nd_real,
                                 const int32_t *d_perm, const sph_sources *src,
                                 const double *d_rec, int uniform_h, int accumulate, double *rho,
                                 double *out_s, int64_t out_s_stride);
int sph_momentum_av_staged(sph_ctx *ctx, const sph_grid *grid, double radius_scale, int dim,
                           int64_t nd_total, int64_t nd_real, const int32_t *d_perm,
                           const sph_sources *src, const double *d_rec, int uniform_h,
                           double alpha, double beta, double eta, int accumulate, double *ax,
                           double *ay, double *az, double *dt_fac);

/* SPHDensityRate (sph/funcs/density_funcs.pyx:116-164, hks off): the
 * continuity equation rate_a = sum_b (v_a - v_b) . gradW_a(r_ab, h_ab) m_b over
 * the SELL-64 lists, in list order; out (original order) = [out +] rate, 0 for
 * non-LocalReal destinations (sph_func.pyx:208-227).  Records need x y z m u
 * v w (and h when !uniform_h). */
int sph_density_rate(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                     const int32_t *d_perm, const int32_t *counts, const int64_t *slice_off,
                     const int32_t *indices, const double *d_rec, const double *s_rec,
                     int uniform_h, int accumulate, double *out);
int sph_density_rate_f32(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                         const int32_t *d_perm, const int32_t *counts, const int64_t *slice_off,
                         const int32_t *indices, const float *d_rec, const float *s_rec,
                         int uniform_h, int accumulate, double *out);

/* XSPHCorrection (sph/funcs/xsph_funcs.pyx:36-91, hks and the Bonnet-Lok
 * correction off): out_c = [out_c +] sum_b (m_b W_ab / rho_ab) (v_b - v_a)_c eps,
 * rho_ab = (rho_b + rho_a)/2; oy, oz may be NULL.  Records need x y z m u v w
 * rho (and h when !uniform_h). */
int sph_xsph(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real, const int32_t *d_perm,
             const int32_t *counts, const int64_t *slice_off, const int32_t *indices,
             const double *d_rec, const double *s_rec, int uniform_h, double eps,
             int accumulate, double *ox, double *oy, double *oz);
int sph_xsph_f32(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                 const int32_t *d_perm, const int32_t *counts, const int64_t *slice_off,
                 const int32_t *indices, const float *d_rec, const float *s_rec, int uniform_h,
                 double eps, int accumulate, double *ox, double *oy, double *oz);

/* Split-form pair operators and the wall force, over the same lists and
 * records as sph_momentum_av (one launch for all source arrays of a calc,
 * summed per source from 0 in list order and added in source order; outputs
 * 0 for non-LocalReal destinations, sph_func.pyx:208-227):
 *   SPH_PAIR_PRESSURE_GRADIENT     SPHPressureGradient.eval_nbr
 *                                  (sph/funcs/pressure_funcs.pyx:61-151)
 *   SPH_PAIR_ARTIFICIAL_VISCOSITY  MonaghanArtificialViscosity.eval_nbr
 *                                  (sph/funcs/viscosity_funcs.pyx:62-151), dt_fac
 *                                  max-reduced with the 0.01 constant of :99
 *   SPH_PAIR_MOMENTUM              MomentumEquation.eval_nbr with the artificial
 *                                  pressure branch (pressure_funcs.pyx:246-391,
 *                                  :327-365), dt_fac max-reduced
 *   SPH_PAIR_BOUNDARY_FORCE        MonaghanBoundaryForce.eval_nbr
 *                                  (sph/funcs/boundary_funcs.pyx:55-107); the
 *                                  sources' tangent/normal frames come from
 *                                  `frames` (original order, addressed through
 *                                  each source's sorted->original permutation)
 * with_correction switches the artificial pressure term on (the functions'
 * deltap argument, pressure_funcs.pyx:27-40, :185-205).  A pair whose
 * reference arithmetic divides by zero (W(deltap) == 0, a wall distance y ==
 * 0, delp == 0, m_a + m_b == 0) raises in the reference's Cython and adds
 * nothing; it adds nothing here.  ns of each source: records whose P/rho^2
 * field is refreshed first (pressure gradient and momentum; 0 when current).
 * dt_fac may be NULL for the pressure gradient and the wall force. */
#define SPH_PAIR_PRESSURE_GRADIENT 1
#define SPH_PAIR_ARTIFICIAL_VISCOSITY 2
#define SPH_PAIR_MOMENTUM 3
#define SPH_PAIR_BOUNDARY_FORCE 4
typedef struct sph_pair_params {
    int32_t op;              /* SPH_PAIR_* */
    int32_t with_correction; /* artificial pressure on (gradient, momentum) */
    double alpha, beta, eta; /* viscosity; the viscosity function's eta is 0
                                unless set after construction (viscosity_funcs.pyx:30-43) */
    double deltap, n, eps;   /* artificial pressure */
    double delp;             /* wall force: reach along the wall */
} sph_pair_params;
typedef struct sph_wall_frames {
    const int32_t *s_perm[SPH_MAX_SRC]; /* sorted slot -> source particle index */
    const double *tx[SPH_MAX_SRC], *ty[SPH_MAX_SRC], *tz[SPH_MAX_SRC];
    const double *nx[SPH_MAX_SRC], *ny[SPH_MAX_SRC], *nz[SPH_MAX_SRC];
} sph_wall_frames;
int sph_pair_accel(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                   const int32_t *d_perm, const sph_sources *src, const double *d_rec,
                   int uniform_h, const sph_pair_params *prm, const sph_wall_frames *frames,
                   int accumulate, double *ax, double *ay, double *az, double *dt_fac);
int sph_pair_accel_f32(sph_ctx *ctx, int dim, int64_t nd_total, int64_t nd_real,
                       const int32_t *d_perm, const sph_sources *src, const float *d_rec,
                       int uniform_h, const sph_pair_params *prm,
                       const sph_wall_frames *frames, int accumulate, double *ax, double *ay,
                       double *az, double *dt_fac);

/* Reductions of the time-step functions (solver/time_step_functions.py:8-57,
 * Integrator.get_max_acceleration, solver/integrator.py:369-412): over n
 * elements, out[0] = max |(a, b, c)| (sqrt(a*a + b*b [+ c*c]) with ncomp
 * components, |a| with one), out[1] = min a, out[2] = max a; NaN propagates as
 * in numpy.  out is host memory; the call waits for the stream. */
int sph_extrema(sph_ctx *ctx, int64_t n, int ncomp, const double *a, const double *b,
                const double *c, double *out);

/* The reductions of sph_extrema without allocating anything.  No counterpart
 * in the reference: for callers that cannot allocate device or pinned memory
 * (code under HIP graph capture, tight stepping loops, harnesses that forbid
 * hipMalloc).  The block partials live in a file-scope device array of the
 * library (3 x 512 doubles, 12 KiB per device, held until the process ends);
 * the three results (max |a|, min a, max a) are written to out_dev, device
 * memory of 3 doubles, stream-ordered: the call neither allocates nor waits.
 * Any n is accepted (at most 512 blocks, grid-stride).  Every call on one
 * device shares that array, so calls that overlap on one device (other
 * streams, other host threads) race. */
int sph_extrema_noalloc(sph_ctx *ctx, int64_t n, int ncomp, const double *a, const double *b,
                        const double *c, double *out_dev);

/* GravityForce (sph/funcs/external_force.pyx:38-44): a (+)= g; az may be
 * NULL.  Original order. */
int sph_gravity(sph_ctx *ctx, int64_t n_real, int64_t n_total, double gx, double gy, double gz,
                int accumulate, double *ax, double *ay, double *az);

/* PositionStepping (sph/funcs/position_funcs.pyx:29-43): a_x = u
 * (assignment).  Original order. */
int sph_position_stepping(sph_ctx *ctx, int64_t n_real, int64_t n_total, const double *u,
                          const double *v, const double *w, double *ax, double *ay, double *az);

/* CubicSplineKernel.function / .gradient (base/kernels.pyx:355-417) for n
 * point pairs: w[i] = W(a_i - b_i, h_i), g*[i] = gradW components (any
 * output may be NULL). */
int sph_cubic_spline(sph_ctx *ctx, int dim, int64_t n, const double *ax, const double *ay,
                     const double *az, const double *bx, const double *by, const double *bz,
                     const double *h, double *w, double *gx, double *gy, double *gz);

/* ---- integrator stage arithmetic (solver/integrator.py) ----------------- */

typedef struct sph_step_props {
    int32_t nprops;
    double *x[SPH_MAX_PROPS];         /* stepped property (out) */
    const double *x0[SPH_MAX_PROPS];  /* initial value */
    const double *a[SPH_MAX_PROPS];   /* acceleration */
} sph_step_props;

/* Integrator.step (integrator.py:340-367): x = x0 + a*dt for every prop. */
int sph_stage_step(sph_ctx *ctx, int64_t n, const sph_step_props *props, double dt);
/* PredictorCorrectorIntegrator.final_step (integrator.py:632-643): x = 2x - x0 */
int sph_pc_final(sph_ctx *ctx, int64_t n, const sph_step_props *props);
/* copy n doubles (save_initial_arrays, integrator.py:282-285) */
int sph_copy(sph_ctx *ctx, int64_t n, const double *src, double *dst);
/* fill n doubles with value (reset_accelerations / set_to_zero) */
int sph_fill(sph_ctx *ctx, int64_t n, double value, double *dst);
/* the same for several arrays in one launch: dst[k][i] = src[k][i], or value
 * (src and dst_stride unused by the fill; the arrays are contiguous) */
int sph_copy_many(sph_ctx *ctx, int64_t n, const sph_prop_list *props);
int sph_fill_many(sph_ctx *ctx, int64_t n, double value, const sph_prop_list *props);

/* ---- batched simulations ------------------------------------------------ */

/* Many independent simulations packed into one set of arrays (simulation s
 * owns original indices [sim_off[s], sim_off[s+1]); sim_of[i] = s), each
 * with its own cell grid (cell size radius_scale*max h of its particles,
 * its own bounding box), so each one's neighbour lists and results are those
 * of the single-simulation path.  No reference counterpart: the reference
 * runs one simulation per Solver (solver/solver.py:523-590); a step here is
 * RK2Integrator.integrate (solver/integrator.py:467-498) over SPHRho,
 * TaitEquation, MomentumEquation (+ artificial viscosity) and
 * PositionStepping, for every simulation at once.
 * All arrays are device arrays of n doubles in the packed original order;
 * z, w are read in 2D too (zeros), z0, w0, az, aw only in 3D.
 * The step never synchronises with the host: capacities bound every size and
 * violations set bits of *status (1: neighbour slots > list_cap, 2: cell keys
 * > key_cap, 4: non-finite coordinate or h, 8: grid too large, 16: non-empty
 * cells > max_cells, 32: reverse mode over particles of unequal h, 64: the
 * persistent evaluation's grid barrier timed out); a step
 * with a nonzero status is void.  *pairs
 * accumulates the list entries of every build (each build's lists serve two
 * operators). */
typedef struct sph_batch_desc {
    int32_t nsims, dim;
    int64_t n;
    const int64_t *sim_off; /* nsims + 1 */
    const int32_t *sim_of;  /* n */
    double *x, *y, *z, *u, *v, *w, *h, *m, *rho, *p, *cs;
    double *x0, *y0, *z0, *u0, *v0, *w0; /* stage initial values */
    double *ax, *ay, *az, *au, *av, *aw; /* accelerations of x, y, z, u, v, w */
    double *dt_fac;
    const double *dt_sim; /* nsims time steps, or NULL; when set, the dt of a
                           * step or stage call multiplies dt_sim[s] */
    double co, ro, gamma, alpha, beta, eta, radius_scale;
    int64_t key_cap;   /* cells of all grids together */
    int64_t max_cells; /* non-empty cells */
    int64_t list_cap;  /* neighbour-list slots */
    int32_t maxch;     /* candidate chunks of 64 replayed by the list fill, 1..16 */
    int32_t f32;       /* 1: float records, lists and operators (state stays FP64) */
    int32_t uniform_h; /* 1: every simulation has one h; sources' h is not read (a
                        * simulation whose h differs sets status bit 32) */
    /* launch shapes of sph_batch_rk2_step (0 picks the default); any value
     * gives the same results, bit for bit (the per-particle sums keep their
     * order) -- an autotuner picks them by timing */
    int32_t op_threads;       /* threads per block of the fused operator kernels:
                               * 64, 128 or 256 (default 256) */
    int32_t op_blocks_per_cu; /* persistent grid of the fused kernels: blocks per
                               * compute unit, each walking a contiguous range of
                               * particle tiles (default 0: one tile per block) */
    int32_t nbr_tile;         /* neighbour records in flight per lane in the fused
                               * momentum kernel: 1, 2 or 4 (density: twice that;
                               * default 2) */
    int32_t cells_per_block;  /* destination cells each block of the list count
                               * pass walks: 1..64 (default 1) */
    void *work;        /* device workspace of sph_batch_workspace_size bytes */
    int64_t work_bytes;
    int32_t *status;   /* device */
    int64_t *pairs;    /* device */
    int32_t persistent;       /* 1: each evaluation after its lists is ONE launch --
                               * SPHRho + Tait, a grid barrier, MomentumEquation +
                               * PositionStepping + stage -- over at most the
                               * blocks the device holds at once (status bit 64 if
                               * the barrier's bounded wait runs out); 0: two
                               * launches.  Same results either way. */
    int32_t staged;           /* 1: the two fused kernels with the neighbour records
                               * staged in LDS pencil by pencil (one wavefront per
                               * 64-row slice; op_threads, op_blocks_per_cu, nbr_tile
                               * and persistent then unused).  Same results. */
} sph_batch_desc;

int sph_batch_workspace_size(sph_ctx *ctx, const sph_batch_desc *d, int64_t *bytes);
/* prepare a new workspace (clears its record area); call once per workspace */
int sph_batch_init_workspace(sph_ctx *ctx, const sph_batch_desc *d);
/* one RK2 step of every simulation, stream-ordered (RK2Integrator.integrate,
 * solver/integrator.py:467-498).  Per evaluation: binning and lists, then two
 * fused kernels -- SPHRho + Tait, and MomentumEquation + PositionStepping +
 * the stage update (with save_initial_arrays folded into the first stage) */
int sph_batch_rk2_step(sph_ctx *ctx, const sph_batch_desc *d, double dt);
/* capture nsteps RK2 steps in a HIP graph (the context needs its own
 * stream; run one eager step first so the context scratch has its final
 * size -- the context then refuses to grow it) */
int sph_batch_graph_create(sph_ctx *ctx, const sph_batch_desc *d, double dt, int nsteps,
                           void **graph);
int sph_batch_graph_launch(sph_ctx *ctx, void *graph);
int sph_batch_graph_destroy(sph_ctx *ctx, void *graph);

/* The pieces of a step, for callers that drive the stages themselves (the
 * reverse mode recomputes them): x0 <- x etc. (Integrator.save_initial_arrays,
 * solver/integrator.py:262-285); one evaluation at the current positions
 * (binning, lists, SPHRho, Tait, MomentumEquation, PositionStepping), which
 * leaves the lists and sorted records in the workspace; x <- x0 + a*dt for
 * x y (z) u v (w) (Integrator.step, :340-367). */
int sph_batch_save_initial(sph_ctx *ctx, const sph_batch_desc *d);
int sph_batch_eval(sph_ctx *ctx, const sph_batch_desc *d);
int sph_batch_stage(sph_ctx *ctx, const sph_batch_desc *d, double dt);

/* Reverse mode of the evaluation last run by sph_batch_eval on the same
 * workspace, over its frozen neighbour lists.  With a = the MomentumEquation
 * accelerations (densities from SPHRho, pressures and sound speeds from
 * Tait) and abar their cotangent, writes xbar = (da/dx)^T abar,
 * ubar = (da/du)^T abar (the dx/dt = u part is the caller's) and
 * *alpha_bar = (da/dalpha)^T abar (device scalar).  Every simulation must
 * have one h for all its particles (then the lists are symmetric and each
 * particle gathers the adjoint of every pair term from its own row, with no
 * atomics); status |= 32 otherwise.  No reference counterpart: the reference
 * has no derivatives. */
int sph_batch_vjp(sph_ctx *ctx, const sph_batch_desc *d, const double *abar_u,
                  const double *abar_v, const double *abar_w, double *xbar, double *ybar,
                  double *zbar, double *ubar, double *vbar, double *wbar, double *alpha_bar);

#ifdef __cplusplus
}
#endif

#endif /* SPHHIP_H */

// Many independent simulations advanced together.
//
// The simulations are packed into one set of particle arrays (simulation s
// owns original indices [sim_off[s], sim_off[s+1])).  Each simulation keeps
// its own cell grid -- cell size radius_scale*max(h) over its particles and
// its own bounding box, exactly what CellManager would build for it alone
// (base/cell.pyx:911-941, :1140-1158) -- so its neighbour lists, and with
// them every sum, are the ones the single-simulation path produces.  The
// grids are laid side by side in one key space (key_base[s] + linear cell
// index), so one binning pass, one list build and one launch per operator
// serve every simulation.
//
// A step runs without any host synchronisation: sizes are bounded by the
// workspace capacities and checked on the device (status bits), so a whole
// step, or several, can be captured in a HIP graph and replayed.
//
// The step is RK2Integrator.integrate (solver/integrator.py:467-498) over
// the operation list of the 2D/3D drop: SPHRho (non-integrating), Tait
// (non-integrating), MomentumEquation with artificial viscosity and
// PositionStepping (integrating).
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sph_internal.h"
#include "sph_staged.h"
#include "sph_terms.h"

#define BT_THREADS 256

// ---------------------------------------------------------------------------
// per-simulation grids
// ---------------------------------------------------------------------------
#define GRID_THREADS 1024  // one block per simulation for its bounding box

__device__ __forceinline__ double blk_min(double v, double *red)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fmin(v, __shfl_xor(v, d, 64));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double r = red[0];
    for (int w = 1; w < GRID_THREADS / 64; ++w) r = fmin(r, red[w]);
    return r;
}

// One block per simulation: bounding box and h range of its particles, then
// its grid (CellManager.compute_cell_size with the Particles defaults: no
// clamps, 0 -> 1.0).  ncell[s] = cells of the grid.
__global__ void __launch_bounds__(GRID_THREADS)
    k_batch_grid(const int64_t *__restrict__ sim_off, const double *__restrict__ x,
                 const double *__restrict__ y, const double *__restrict__ z,
                 const double *__restrict__ h, double radius_scale, int uniform_h,
                 GridDev *__restrict__ grids, int64_t *__restrict__ ncell,
                 int32_t *__restrict__ status)
{
    __shared__ double red[GRID_THREADS / 64];
    const int s = blockIdx.x;
    const int64_t b = sim_off[s], e = sim_off[s + 1];
    double v[8];
    for (int k = 0; k < 8; ++k) v[k] = INFINITY;
    double bad = 0.0;
    for (int64_t i = b + threadIdx.x; i < e; i += GRID_THREADS) {
        const double a = x[i], bb = y[i], c = z[i], hh = h[i];
        if (!isfinite(a) || !isfinite(bb) || !isfinite(c) || !isfinite(hh)) {
            bad = 1.0;
            continue;
        }
        v[0] = fmin(v[0], a);
        v[1] = fmin(v[1], bb);
        v[2] = fmin(v[2], c);
        v[3] = fmin(v[3], -a);
        v[4] = fmin(v[4], -bb);
        v[5] = fmin(v[5], -c);
        v[6] = fmin(v[6], -hh);
        v[7] = fmin(v[7], hh);
    }
    for (int k = 0; k < 8; ++k) v[k] = blk_min(v[k], red);
    bad = -blk_min(-bad, red);
    if (threadIdx.x != 0) return;
    GridDev g;
    g.tab = nullptr;
    g.key_base = 0;
    int64_t cells = 1;
    if (e <= b || bad > 0.0) {
        if (bad > 0.0) atomicOr(status, 4);
        g.cs = 1.0;
        g.cmin0 = g.cmin1 = g.cmin2 = 0;
        g.d0 = bad > 0.0 ? 0 : 1;  // d0 = 0: void, its particles go to the trash cell
        g.d1 = g.d2 = 1;
    } else {
        if (uniform_h && -v[6] != v[7]) atomicOr(status, 32);  // h differs within the simulation
        double cs = radius_scale * (-v[6]);
        if (cs == 0.0) cs = 1.0;
        g.cs = cs;
        int cm[3], dm[3];
        for (int a = 0; a < 3; ++a) {
            const double fl = floor(v[a] / cs), fh = floor(-v[a + 3] / cs);
            if (fabs(fl) > 1.0e9 || fabs(fh) > 1.0e9 || fh - fl + 1.0 > 1.0e6) {
                atomicOr(status, 8);
                cm[a] = 0;
                dm[a] = 0;  // void (see above)
            } else {
                cm[a] = (int)fl;
                dm[a] = (int)(fh - fl) + 1;
            }
            cells *= dm[a];
        }
        g.cmin0 = cm[0];
        g.cmin1 = cm[1];
        g.cmin2 = cm[2];
        g.d0 = (dm[0] && dm[1] && dm[2]) ? dm[0] : 0;
        g.d1 = dm[1] ? dm[1] : 1;
        g.d2 = dm[2] ? dm[2] : 1;
    }
    grids[s] = g;
    ncell[s] = g.d0 ? cells : 0;
}

// key_base of every simulation (exclusive sum of the cell counts).  The
// last key, key_cap - 1, is the trash cell: particles of a void simulation
// (non-finite values, a grid too large, or no room left in the key space --
// status |= 2) go there and get no neighbour lists, so a simulation that
// blew up costs no more than one that did not.
__global__ void __launch_bounds__(1024)
    k_batch_key_base(int nsims, const int64_t *__restrict__ ncell, int64_t key_cap,
                     GridDev *__restrict__ grids, int32_t *__restrict__ status)
{
    // block-wide exclusive scan of the cells of the live simulations, 1024
    // simulations at a time; a simulation that does not fit becomes void
    __shared__ int64_t part[1024 / 64];
    __shared__ int64_t carry;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int s0 = 0; s0 < nsims; s0 += 1024) {
        const int s = s0 + threadIdx.x;
        const bool live = s < nsims && grids[s].d0 != 0;
        const int64_t c = live ? ncell[s] : 0;
        int64_t incl = c;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int64_t t = __shfl_up(incl, d, 64);
            if (lane >= d) incl += t;
        }
        if (lane == 63) part[wid] = incl;
        __syncthreads();
        int64_t before = carry;
        for (int w = 0; w < wid; ++w) before += part[w];
        const int64_t base = before + incl - c;
        if (live) {
            if (base + c <= key_cap - 1) {
                grids[s].key_base = (int32_t)base;
            } else {
                GridDev g = grids[s];
                g.key_base = 0;
                g.d0 = 0;
                grids[s] = g;
                atomicOr(status, 2);
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int64_t tot = 0;
            for (int w = 0; w < 1024 / 64; ++w) tot += part[w];
            carry += tot;
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(BT_THREADS)
    k_batch_keys(const GridDev *__restrict__ grids, const int32_t *__restrict__ sim_of,
                 const double *__restrict__ x, const double *__restrict__ y,
                 const double *__restrict__ z, int64_t n, int64_t key_cap,
                 int32_t *__restrict__ count, int32_t *__restrict__ cell_of,
                 int32_t *__restrict__ slot, int32_t *__restrict__ key_sim,
                 int32_t *__restrict__ status)
{
    const int64_t i = (int64_t)blockIdx.x * BT_THREADS + threadIdx.x;
    const bool act = i < n;  // the active lanes are a prefix of the wavefront
    int32_t key32 = -1, sim = 0;
    if (act) {
        sim = sim_of[i];
        const GridDev g = grids[sim];
        int64_t key = key_cap - 1;  // trash cell of void simulations
        if (g.d0 > 0) {
            const int cx = min(max(cell_coord(x[i], g.cs) - g.cmin0, 0), g.d0 - 1);
            const int cy = min(max(cell_coord(y[i], g.cs) - g.cmin1, 0), g.d1 - 1);
            const int cz = min(max(cell_coord(z[i], g.cs) - g.cmin2, 0), g.d2 - 1);
            key = (int64_t)g.key_base + ((int64_t)cx * g.d1 + cy) * g.d2 + cz;
            if (key >= key_cap - 1) {  // cannot happen after k_batch_key_base; kept in bounds
                atomicOr(status, 2);
                key = key_cap - 1;
            }
        }
        key32 = (int32_t)key;
        cell_of[i] = key32;
    }
    // one atomic per run of consecutive lanes in the same cell (slots need
    // only be distinct within a cell)
    const int lane = __lane_id();
    const int32_t prev = __shfl_up(key32, 1, 64);
    const bool head = act && (lane == 0 || prev != key32);
    const uint64_t heads = __builtin_amdgcn_ballot_w64(head);
    const uint64_t acts = __builtin_amdgcn_ballot_w64(act);
    const uint64_t le = heads & ((2ull << lane) - 1ull);
    const int start = le ? 63 - __clzll(le) : 0;
    const uint64_t above = heads & ~((2ull << start) - 1ull);
    const int next = above ? __ffsll((long long)above) - 1 : __popcll(acts);
    int32_t base = 0;
    if (head) {
        base = atomicAdd(&count[key32], next - start);
        // every particle of a key is of one simulation (the keys of a
        // simulation are its own range; the trash key's is never read)
        key_sim[key32] = sim;
    }
    base = __shfl(base, start, 64);
    if (act) slot[i] = base + (lane - start);
}

// more non-empty cells than the list kernels are launched for: status |= 16
__global__ void k_batch_check_cells(const int32_t *__restrict__ ncell_dev, int64_t max_cells,
                                    int32_t *__restrict__ status)
{
    if (threadIdx.x == 0 && *ncell_dev > max_cells) *status |= 16;
}

// x = x0 + a * (scale * dt_sim[s]) for the stepped properties (the
// per-simulation time step of Integrator.step, solver/integrator.py:340-367)
__global__ void __launch_bounds__(BT_THREADS)
    k_batch_stage(int64_t n, const int32_t *__restrict__ sim_of, const double *__restrict__ dt_sim,
                  double scale, sph_step_props pr)
{
    const int64_t i = (int64_t)blockIdx.x * BT_THREADS + threadIdx.x;
    if (i >= n) return;
    const double dt = scale * dt_sim[sim_of[i]];
    for (int k = 0; k < pr.nprops; ++k) {
        const double t = pr.a[k][i] * dt;
        pr.x[k][i] = pr.x0[k][i] + t;
    }
}

// ---------------------------------------------------------------------------
// fused kernels of sph_batch_rk2_step
// ---------------------------------------------------------------------------
// Two kernels per evaluation once the lists are built:
//   k_batch_density_eos    SPHRho, then Tait and P/rho^2 of the same particle
//   k_batch_momentum_stage MomentumEquation + AV, then PositionStepping and
//                          x <- x0 + a*dt of the same particle (the first
//                          stage of a step also does save_initial_arrays)
// Each keeps the per-pair order and the arithmetic of the separate kernels
// (k_summation_density, k_tait_rec, k_momentum, k_position_stepping,
// k_batch_stage), so the results are the same bit for bit.  The density
// kernel writes fields 7, 8, 9 and 11 of the records while other lanes read
// fields 0-3 and 10 of their neighbours: distinct words.  The momentum
// kernel reads only records and writes the original-order state.
//
// Launch shape (tunable): blockDim.x threads per block (64, 128 or 256; a
// multiple of the 64-row slices of the lists); a persistent grid walks the
// tiles of blockDim.x sorted particles, block b taking a contiguous range
// of tiles after the XCD remap, so each XCD's blocks cover one contiguous
// part of the sorted order.  NT neighbour records are loaded before their
// terms are added in list order.
#define FUSED_MAX_THREADS 256

__device__ __forceinline__ void fused_tiles(int64_t ntiles, int64_t &tb, int64_t &te)
{
    const int64_t G = gridDim.x, L = sph_xcd_tile(blockIdx.x, gridDim.x);
    tb = L * ntiles / G;
    te = (L + 1) * ntiles / G;
}

// the density sum nr of sorted particle t (original index i) as SPHRho's
// output, then Tait and P/rho^2 of it, into the state and the record
template <typename T>
__device__ __forceinline__ void eos_epilogue(int64_t t, int32_t i, T nr, T *__restrict__ rec, T ro,
                                             T co, T gamma, T B, double *__restrict__ rho,
                                             double *__restrict__ p, double *__restrict__ cs)
{
    // SPHRho is non-integrating: the output starts from 0 (sph_calc.pyx:319-321)
    const T rhoa = T(0.0) + nr;
    // Tait (eos_funcs.pyx:118-129), as k_tait_rec
    const T ratio = rhoa / ro;
    const T gamma2 = T(0.5) * (gamma - T(1.0));
    const T tmp = sph_pow(ratio, gamma);
    const T r0 = T(0.0) + (tmp - T(1.0)) * B;
    const T r1 = T(0.0) + sph_pow(ratio, gamma2) * co;
    const T rhob2 = T(1.0) / (rhoa * rhoa);
    rho[i] = rhoa;
    p[i] = r0;
    cs[i] = r1;
    T *rw = rec + t * SPH_REC;
    rw[SPH_R_RHO] = rhoa;
    rw[SPH_R_P] = r0;
    rw[SPH_R_CS] = r1;
    rw[SPH_R_PRHO2] = r0 * rhob2;
}

// SPHRho, then Tait and P/rho^2, of sorted particle t
template <bool UH, typename T, int NT>
__device__ __forceinline__ void density_eos_one(int dim, int64_t t,
                                                const int32_t *__restrict__ perm,
                                                const int32_t *__restrict__ counts,
                                                const int64_t *__restrict__ slice_off,
                                                const int32_t *__restrict__ indices,
                                                T *__restrict__ rec, T ro, T co, T gamma, T B,
                                                double *__restrict__ rho, double *__restrict__ p,
                                                double *__restrict__ cs)
{
    const T *ra = rec + t * SPH_REC;
    const auto *dr = rec2<T>(rec, t);
    const auto dxy = dr[0], dzm = dr[1];
    const T xa = dxy.x, ya = dxy.y, za = dzm.x, ha = ra[SPH_R_H];
    const T facA = sph_h_dim(ha, dim) * sph_cubic_fac1<T>(dim);
    const T invA = sph_inv_for_div(ha);
    T nr = T(0.0);
    const int32_t *row = indices + slice_off[t >> 6] + (t & 63);
    const int32_t e = counts[t];
    int32_t k = 0;
    for (; k + NT <= e; k += NT) {
        int32_t j[NT];
        typename Pair<T>::type a[NT], b[NT];
        T hb[NT];
#pragma unroll
        for (int u = 0; u < NT; ++u)
            j[u] = __builtin_nontemporal_load(row + ((int64_t)(k + u) << 6));
#pragma unroll
        for (int u = 0; u < NT; ++u) {
            const auto *sr = rec2<T>(rec, j[u]);
            a[u] = sr[0];
            b[u] = sr[1];
            hb[u] = UH ? ha : rec[(int64_t)j[u] * SPH_REC + SPH_R_H];
        }
#pragma unroll
        for (int u = 0; u < NT; ++u)
            nr += rho_term<UH, T>(dim, xa, ya, za, ha, facA, a[u], b[u], hb[u], invA);
    }
    for (; k < e; ++k) {
        const int32_t j = __builtin_nontemporal_load(row + ((int64_t)k << 6));
        const auto *sr = rec2<T>(rec, j);
        const T hb = UH ? ha : rec[(int64_t)j * SPH_REC + SPH_R_H];
        nr += rho_term<UH, T>(dim, xa, ya, za, ha, facA, sr[0], sr[1], hb, invA);
    }
    eos_epilogue<T>(t, perm[t], nr, rec, ro, co, gamma, B, rho, p, cs);
}

template <bool UH, typename T, int NT>
__global__ void __launch_bounds__(FUSED_MAX_THREADS)
    k_batch_density_eos(int dim, int64_t n, int64_t ntiles, const int32_t *__restrict__ perm,
                        const int32_t *__restrict__ counts, const int64_t *__restrict__ slice_off,
                        const int32_t *__restrict__ indices, T *__restrict__ rec, T ro, T co,
                        T gamma, T B, double *__restrict__ rho, double *__restrict__ p,
                        double *__restrict__ cs)
{
    int64_t tb, te;
    fused_tiles(ntiles, tb, te);
    for (int64_t tile = tb; tile < te; ++tile) {
        const int64_t t = tile * blockDim.x + threadIdx.x;
        if (t >= n) break;
        density_eos_one<UH, T, NT>(dim, t, perm, counts, slice_off, indices, rec, ro, co, gamma,
                                   B, rho, p, cs);
    }
}

struct FusedState {
    double *x[3], *u[3], *x0[3], *u0[3], *ax[3], *au[3];
    double *dt_fac;
};

// dt_fac, the momentum sums as the accelerations of u, v, w (integrating
// calc: output 0 + sum), PositionStepping's a_x = u, then x <- x0 + a*dt per
// stepped property (x0 <- x first with save), of original particle i
template <typename T>
__device__ __forceinline__ void stage_epilogue(int dim, int32_t i, T dtf, T nr0, T nr1, T nr2,
                                               const FusedState &st,
                                               const int32_t *__restrict__ sim_of,
                                               const double *__restrict__ dt_sim, double scale,
                                               int save)
{
    st.dt_fac[i] = dtf;
    const double acc[3] = {(double)(T(0.0) + nr0), (double)(T(0.0) + nr1),
                           (double)(T(0.0) + nr2)};
    const double dt = dt_sim ? scale * dt_sim[sim_of[i]] : scale;
    // every component is read before the first store (memory operations
    // retire in order: a load behind a store waits for it)
    double uc[3] = {0.0, 0.0, 0.0}, xc[3] = {0.0, 0.0, 0.0};
    double u0[3] = {0.0, 0.0, 0.0}, x0[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (c >= dim) break;
        uc[c] = st.u[c][i];
        xc[c] = st.x[c][i];
        u0[c] = save ? uc[c] : st.u0[c][i];
        x0[c] = save ? xc[c] : st.x0[c][i];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (c >= dim) break;
        if (save) {
            st.u0[c][i] = uc[c];
            st.x0[c][i] = xc[c];
        }
        st.au[c][i] = acc[c];
        st.ax[c][i] = uc[c];
        const double tu = acc[c] * dt, tx = uc[c] * dt;
        st.u[c][i] = u0[c] + tu;
        st.x[c][i] = x0[c] + tx;
    }
}

// MomentumEquation + AV, then PositionStepping and the stage update, of
// sorted particle t
template <bool UH, typename T, int NT>
__device__ __forceinline__ void momentum_stage_one(int dim, int64_t t,
                                                   const int32_t *__restrict__ perm,
                                                   const int32_t *__restrict__ counts,
                                                   const int64_t *__restrict__ slice_off,
                                                   const int32_t *__restrict__ indices,
                                                   const T *__restrict__ rec, T alpha, T beta,
                                                   T eta, const FusedState &st,
                                                   const int32_t *__restrict__ sim_of,
                                                   const double *__restrict__ dt_sim,
                                                   double scale, int save)
{
    const auto *dr = rec2<T>(rec, t);
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
    T dtf = T(0.0);  // dt_fac is reset per calc (pressure_funcs.pyx:214)
    T nr0 = T(0.0), nr1 = T(0.0), nr2 = T(0.0);
    const int32_t *row = indices + slice_off[t >> 6] + (t & 63);
    const int32_t e = counts[t];
    int32_t k = 0;
    for (; k + NT <= e; k += NT) {
        int32_t j[NT];
        typename Pair<T>::type s0[NT], s1[NT], s2[NT], s3[NT], s4[NT];
        T hb[NT];
#pragma unroll
        for (int u = 0; u < NT; ++u)
            j[u] = __builtin_nontemporal_load(row + ((int64_t)(k + u) << 6));
#pragma unroll
        for (int u = 0; u < NT; ++u) {
            const auto *sr = rec2<T>(rec, j[u]);
            s0[u] = sr[0];
            s1[u] = sr[1];
            s2[u] = sr[2];
            s3[u] = sr[3];
            s4[u] = sr[4];
            hb[u] = UH ? D.ha : rec[(int64_t)j[u] * SPH_REC + SPH_R_H];
        }
#pragma unroll
        for (int u = 0; u < NT; ++u)
            mom_term<UH, T>(dim, D, alpha, beta, eta, s0[u], s1[u], s2[u], s3[u], s4[u],
                            hb[u], dtf, nr0, nr1, nr2);
    }
    for (; k < e; ++k) {
        const int32_t j = __builtin_nontemporal_load(row + ((int64_t)k << 6));
        const auto *sr = rec2<T>(rec, j);
        const T hb = UH ? D.ha : rec[(int64_t)j * SPH_REC + SPH_R_H];
        mom_term<UH, T>(dim, D, alpha, beta, eta, sr[0], sr[1], sr[2], sr[3], sr[4], hb, dtf,
                        nr0, nr1, nr2);
    }
    stage_epilogue<T>(dim, perm[t], dtf, nr0, nr1, nr2, st, sim_of, dt_sim, scale, save);
}

template <bool UH, typename T, int NT>
__global__ void __launch_bounds__(FUSED_MAX_THREADS)
    k_batch_momentum_stage(int dim, int64_t n, int64_t ntiles, const int32_t *__restrict__ perm,
                           const int32_t *__restrict__ counts,
                           const int64_t *__restrict__ slice_off,
                           const int32_t *__restrict__ indices, const T *__restrict__ rec,
                           T alpha, T beta, T eta, FusedState st,
                           const int32_t *__restrict__ sim_of, const double *__restrict__ dt_sim,
                           double scale, int save)
{
    int64_t tb, te;
    fused_tiles(ntiles, tb, te);
    for (int64_t tile = tb; tile < te; ++tile) {
        const int64_t t = tile * blockDim.x + threadIdx.x;
        if (t >= n) break;
        momentum_stage_one<UH, T, NT>(dim, t, perm, counts, slice_off, indices, rec, alpha, beta,
                                      eta, st, sim_of, dt_sim, scale, save);
    }
}

// ---------------------------------------------------------------------------
// one evaluation and its stage as one persistent launch
// ---------------------------------------------------------------------------
// The grid is at most the blocks the device holds at once (the occupancy of
// this kernel times the compute units), so every block is resident: the
// blocks walk their tiles' SPHRho + Tait, meet at a grid barrier, then walk
// the same tiles' MomentumEquation + PositionStepping + stage.  The barrier
// publishes the density results with device-scope release and reads them
// after a device-scope acquire (what a kernel boundary does).  Its wait is
// bounded (spin_limit polls of the arrival counter); past the bound -- only
// possible if another kernel kept some blocks from starting -- status |=
// FUSED_WAIT_STATUS, the blocks go on, and the step is reported void.
#define FUSED_WAIT_STATUS 64
// polls of the arrival counter before a block gives up (each poll is a
// device-scope load plus a short sleep, about a microsecond: ~1 s, hundreds of
// times the density phase of the largest batch)
#define FUSED_SPIN_LIMIT ((int64_t)1 << 20)

__device__ __forceinline__ bool fused_grid_barrier(int32_t *arrived, int32_t nblocks,
                                                   int64_t spin_limit)
{
    __shared__ int32_t s_ok;
    __threadfence();  // release: this thread's results, device-wide
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(arrived, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int64_t spins = 0;
        int32_t ok = 1;
        while (__hip_atomic_load(arrived, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < nblocks) {
            if (++spins > spin_limit) {
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        s_ok = ok;
    }
    __syncthreads();
    __threadfence();  // acquire: the other blocks' results, fresh
    return s_ok != 0;
}

template <bool UH, typename T, int NT>
__global__ void __launch_bounds__(FUSED_MAX_THREADS)
    k_batch_fused_eval(int dim, int64_t n, int64_t ntiles, const int32_t *__restrict__ perm,
                       const int32_t *__restrict__ counts, const int64_t *__restrict__ slice_off,
                       const int32_t *__restrict__ indices, T *__restrict__ rec, T ro, T co,
                       T gamma, T B, double *__restrict__ rho, double *__restrict__ p,
                       double *__restrict__ cs, T alpha, T beta, T eta, FusedState st,
                       const int32_t *__restrict__ sim_of, const double *__restrict__ dt_sim,
                       double scale, int save, int32_t *__restrict__ arrived,
                       int32_t *__restrict__ status, int64_t spin_limit)
{
    int64_t tb, te;
    fused_tiles(ntiles, tb, te);
    for (int64_t tile = tb; tile < te; ++tile) {
        const int64_t t = tile * blockDim.x + threadIdx.x;
        if (t < n)
            density_eos_one<UH, T, 2 * NT>(dim, t, perm, counts, slice_off, indices, rec, ro, co,
                                           gamma, B, rho, p, cs);
    }
    if (!fused_grid_barrier(arrived, (int32_t)gridDim.x, spin_limit) && threadIdx.x == 0)
        atomicOr(status, FUSED_WAIT_STATUS);
    for (int64_t tile = tb; tile < te; ++tile) {
        const int64_t t = tile * blockDim.x + threadIdx.x;
        if (t < n)
            momentum_stage_one<UH, T, NT>(dim, t, perm, counts, slice_off, indices, rec, alpha,
                                          beta, eta, st, sim_of, dt_sim, scale, save);
    }
}

// ---------------------------------------------------------------------------
// the two fused kernels with the neighbour records staged in LDS
// ---------------------------------------------------------------------------
// The per-simulation grids key their cells linearly (key_base + (cx*d1 +
// cy)*d2 + cz) and the simulations follow each other in key order, so the
// pencil walk of the staged plain operators (sph_staged.h) applies: one
// wavefront per 64-row slice, each lane with its own simulation's grid.  The
// sums, and the epilogues, are those of k_batch_density_eos and
// k_batch_momentum_stage.
#ifndef BATCH_STG_DENS_CAP
#define BATCH_STG_DENS_CAP 256
#endif
#ifndef BATCH_STG_MOM_CAP
#define BATCH_STG_MOM_CAP 192
#endif

// the lane's grid, cell and reach (records of T; ox, oy, oz: the double
// coordinates the binning used, for float records)
template <typename T>
__device__ __forceinline__ void batch_stage_cell(const GridDev &g, int32_t i, T xa, T ya, T za,
                                                 T ha, const double *__restrict__ ox,
                                                 const double *__restrict__ oy,
                                                 const double *__restrict__ oz, double rs,
                                                 int &cx, int &cy, int &cz, int &dl)
{
    if (ox) stg_cell(g, ox[i], oy[i], oz[i], cx, cy, cz);
    else stg_cell(g, (double)xa, (double)ya, (double)za, cx, cy, cz);
    dl = stg_reach<T>(ha, rs, g.cs);
}

template <bool UH, typename T>
__global__ void __launch_bounds__(STG_THREADS)
    k_batch_density_eos_staged(int dim, int64_t n, const GridDev *__restrict__ grids,
                               const int32_t *__restrict__ sim_of,
                               const int32_t *__restrict__ perm,
                               const int32_t *__restrict__ cell_start,
                               const int32_t *__restrict__ counts,
                               const int64_t *__restrict__ slice_off,
                               const int32_t *__restrict__ indices, T *__restrict__ rec,
                               const double *__restrict__ ox, const double *__restrict__ oy,
                               const double *__restrict__ oz, double rs, T ro, T co, T gamma,
                               T B, double *__restrict__ rho, double *__restrict__ p,
                               double *__restrict__ cs)
{
    using P = typename Pair<T>::type;
    constexpr int NP = UH ? 2 : 3;  // (x, y) (z, m) [(h, -)]
    __shared__ P lds[BATCH_STG_DENS_CAP * NP];
    const int lane = threadIdx.x;
    const int64_t slice = sph_xcd_tile(blockIdx.x, gridDim.x);
    const int64_t t = slice * STG_THREADS + lane;
    const bool act = t < n;
    const int32_t i = act ? perm[t] : 0;
    const int32_t e = act ? counts[t] : 0;
    GridDev g = {};
    T xa = T(0), ya = T(0), za = T(0), ha = T(0), facA = T(0), invA = T(0);
    int cx = 0, cy = 0, cz = 0, dl = 0;
    if (act) {
        const auto *dr = rec2<T>(rec, t);
        const auto dxy = dr[0], dzm = dr[1];
        xa = dxy.x;
        ya = dxy.y;
        za = dzm.x;
        ha = rec[t * SPH_REC + SPH_R_H];
        facA = sph_h_dim(ha, dim) * sph_cubic_fac1<T>(dim);
        invA = sph_inv_for_div(ha);
        if (e > 0) {
            g = grids[sim_of[i]];
            batch_stage_cell<T>(g, i, xa, ya, za, ha, ox, oy, oz, rs, cx, cy, cz, dl);
        }
    }
    const int dmax = stg_wave_max(dl);
    const int32_t *row = indices + slice_off[slice] + lane;
    T nr = T(0.0);
    auto stage = [&](P *dst, int32_t q) {
        const P *sp = rec2<T>(rec, q);
        dst[0] = sp[0];
        dst[1] = sp[1];
        if (!UH) dst[2] = P{rec[(int64_t)q * SPH_REC + SPH_R_H], T(0)};
    };
    auto term = [&](const P *pr) {
        nr += rho_term<UH, T>(dim, xa, ya, za, ha, facA, pr[0], pr[1], UH ? ha : pr[NP - 1].x, invA);
    };
    stg_row<T, NP, BATCH_STG_DENS_CAP>(g, cx, cy, cz, dmax, cell_start, row, e, (int32_t)t, true,
                                       lds, stage, stage, term);
    if (act) eos_epilogue<T>(t, i, nr, rec, ro, co, gamma, B, rho, p, cs);
}

template <bool UH, typename T>
__global__ void __launch_bounds__(STG_THREADS)
    k_batch_momentum_stage_staged(int dim, int64_t n, const GridDev *__restrict__ grids,
                                  const int32_t *__restrict__ sim_of,
                                  const int32_t *__restrict__ perm,
                                  const int32_t *__restrict__ cell_start,
                                  const int32_t *__restrict__ counts,
                                  const int64_t *__restrict__ slice_off,
                                  const int32_t *__restrict__ indices, const T *__restrict__ rec,
                                  const double *__restrict__ ox, const double *__restrict__ oy,
                                  const double *__restrict__ oz, double rs, T alpha, T beta,
                                  T eta, FusedState st, const double *__restrict__ dt_sim,
                                  double scale, int save)
{
    using P = typename Pair<T>::type;
    constexpr int NP = UH ? 5 : 6;  // (x, y) (z, m) (u, v) (w, P/rho^2) (rho, cs) [(h, p)]
    __shared__ P lds[BATCH_STG_MOM_CAP * NP];
    const int lane = threadIdx.x;
    const int64_t slice = sph_xcd_tile(blockIdx.x, gridDim.x);
    const int64_t t = slice * STG_THREADS + lane;
    const bool act = t < n;
    const int32_t i = act ? perm[t] : 0;
    const int32_t e = act ? counts[t] : 0;
    GridDev g = {};
    MomDest<T> D;
    memset(&D, 0, sizeof(D));
    int cx = 0, cy = 0, cz = 0, dl = 0;
    if (act) {
        const auto *dr = rec2<T>(rec, t);
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
        if (e > 0) {
            g = grids[sim_of[i]];
            batch_stage_cell<T>(g, i, D.xa, D.ya, D.za, D.ha, ox, oy, oz, rs, cx, cy, cz, dl);
        }
    }
    const int dmax = stg_wave_max(dl);
    const int32_t *row = indices + slice_off[slice] + lane;
    T dtf = T(0.0);  // dt_fac is reset per calc (pressure_funcs.pyx:214)
    T nr0 = T(0.0), nr1 = T(0.0), nr2 = T(0.0);
    auto stage = [&](P *dst, int32_t q) {
        const P *sp = rec2<T>(rec, q);
#pragma unroll
        for (int f = 0; f < NP; ++f) dst[f] = sp[f];
    };
    auto term = [&](const P *pr) {
        mom_term<UH, T>(dim, D, alpha, beta, eta, pr[0], pr[1], pr[2], pr[3], pr[4],
                        UH ? D.ha : pr[NP - 1].x, dtf, nr0, nr1, nr2);
    };
    stg_row<T, NP, BATCH_STG_MOM_CAP>(g, cx, cy, cz, dmax, cell_start, row, e, (int32_t)t, true,
                                      lds, stage, stage, term);
    if (act) stage_epilogue<T>(dim, i, dtf, nr0, nr1, nr2, st, sim_of, dt_sim, scale, save);
}

// ---------------------------------------------------------------------------
// reverse mode of one evaluation over its (frozen) neighbour lists
// ---------------------------------------------------------------------------
// F(X) = (u, a(x, u)) with a = MomentumEquation + AV over the densities of
// SPHRho and the Tait pressures.  Given the cotangent abar of a, the
// vector-Jacobian product gives xbar = (da/dx)^T abar, ubar = (da/du)^T abar
// and alphabar = (da/dalpha)^T abar.  With one h per simulation the lists
// are symmetric (b in L(a) <=> a in L(b)), so every particle gathers the
// adjoint of each pair term from its own row -- the term of its own
// acceleration (cotangent abar_a) and the term of its neighbour's
// (cotangent abar_b) -- and no atomics are needed.

// gradient factor G (gradW_ab = r_ab G) of the cubic spline and dG/d|r|
__device__ __forceinline__ void cubic_g_dg(double rab, double h, double fac, double &G,
                                           double &dG)
{
    const double q = rab / h;
    if (q > 2.0 || !(q >= 1.0 || q > 1e-14)) {
        G = 0.0;
        dG = 0.0;
    } else if (q >= 1.0) {
        const double t = 2.0 - q;
        G = fac * (-0.75 * t * t / (h * rab));
        dG = fac * 0.75 * t * (2.0 / (h * h * rab) + t / (h * rab * rab));
    } else {
        G = fac * 3.0 * (0.75 * q - 1.0) / (h * h);
        dG = fac * 2.25 / (h * h * h);
    }
}

__global__ void __launch_bounds__(BT_THREADS)
    k_gather3(int64_t n, const int32_t *__restrict__ perm, const double *__restrict__ a0,
              const double *__restrict__ a1, const double *__restrict__ a2,
              double *__restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * BT_THREADS + threadIdx.x;
    if (t >= n) return;
    const int32_t i = perm[t];
    out[3 * t] = a0[i];
    out[3 * t + 1] = a1[i];
    out[3 * t + 2] = a2 ? a2[i] : 0.0;
}

// momentum part: per sorted particle a, xbar/ubar from the pair terms, the
// density adjoint rhobar_a (through P = p/rho^2, the Tait p(rho) and
// cs(rho) and the viscosity's rho and cs) and the alpha terms of its own row
__global__ void __launch_bounds__(BT_THREADS)
    k_vjp_momentum(int64_t n, const int32_t *__restrict__ counts,
                   const int64_t *__restrict__ slice_off, const int32_t *__restrict__ indices,
                   const double *__restrict__ rec, const double *__restrict__ abar, int dim,
                   double alpha, double beta, double eta, double ro, double co, double gamma,
                   double *__restrict__ xbar, double *__restrict__ ubar,
                   double *__restrict__ rhobar, double *__restrict__ alphabar,
                   int32_t *__restrict__ status)
{
    const int64_t t = (int64_t)sph_xcd_tile(blockIdx.x, gridDim.x) * BT_THREADS + threadIdx.x;
    if (t >= n) return;
    const double *ra = rec + t * SPH_REC;
    const double xa = ra[SPH_R_X], ya = ra[SPH_R_Y], za = ra[SPH_R_Z], ma = ra[SPH_R_M];
    const double ua = ra[SPH_R_U], va = ra[SPH_R_V], wa = ra[SPH_R_W];
    const double rhoa = ra[SPH_R_RHO], ca = ra[SPH_R_CS], ha = ra[SPH_R_H], pa = ra[SPH_R_P];
    const double Pa = ra[SPH_R_PRHO2];
    const double fac = sph_h_dim(ha, dim) * sph_cubic_fac1(dim);
    const double Ax = abar[3 * t], Ay = abar[3 * t + 1], Az = abar[3 * t + 2];
    double xb0 = 0, xb1 = 0, xb2 = 0, ub0 = 0, ub1 = 0, ub2 = 0;
    double Pbar = 0, cbar_a = 0, rhobar_pi = 0, albar = 0;
    const int32_t *row = indices + slice_off[t >> 6] + (t & 63);
    const int32_t e = counts[t];
    bool bad_h = false;
    for (int32_t k = 0; k < e; ++k) {
        const int32_t b = row[(int64_t)k << 6];
        if (b == (int32_t)t) continue;  // r = 0: no gradient, no viscosity
        const double *rb = rec + (int64_t)b * SPH_REC;
        if (rb[SPH_R_H] != ha) bad_h = true;
        const double rx = xa - rb[SPH_R_X], ry = ya - rb[SPH_R_Y], rz = za - rb[SPH_R_Z];
        const double vx = ua - rb[SPH_R_U], vy = va - rb[SPH_R_V], vz = wa - rb[SPH_R_W];
        const double mb = rb[SPH_R_M], rhob = rb[SPH_R_RHO], cb = rb[SPH_R_CS];
        const double Pb = rb[SPH_R_PRHO2];
        const double Bx = abar[3 * (int64_t)b], By = abar[3 * (int64_t)b + 1],
                     Bz = abar[3 * (int64_t)b + 2];
        const double r2 = rx * rx + ry * ry + rz * rz;
        const double rab = sqrt(r2);
        double G, dG;
        cubic_g_dg(rab, ha, fac, G, dG);
        const double dot = vx * rx + vy * ry + vz * rz;
        const double den = r2 + eta * eta * ha * ha;
        const double mu = ha * dot / den;
        const double cm = 0.5 * (ca + cb), rm = 0.5 * (rhoa + rhob);
        const bool visc = dot < 0;
        const double Pi = visc ? (-alpha * cm * mu + beta * mu * mu) / rm : 0.0;
        const double s_ab = -mb * (Pa + Pb + Pi), s_ba = -ma * (Pb + Pa + Pi);
        // own term T_ab = s_ab r G (cotangent A); the neighbour's T_ba = -s_ba r G (cotangent B)
        const double gA = (Ax * rx + Ay * ry + Az * rz) * G;
        const double gB = -(Bx * rx + By * ry + Bz * rz) * G;
        Pbar += -mb * gA - ma * gB;
        const double pibar_own = -mb * gA, pibar_tot = pibar_own - ma * gB;
        // gradient: d(r G)/dr = G I + (dG/|r|) r r^T applied to s_ab A (own
        // term) and, with r_ba = -r, to -s_ba B (the neighbour's term)
        const double qx = s_ab * Ax - s_ba * Bx, qy = s_ab * Ay - s_ba * By,
                     qz = s_ab * Az - s_ba * Bz;
        const double rq = (rab > 0.0 ? dG / rab : 0.0) * (rx * qx + ry * qy + rz * qz);
        xb0 += G * qx + rq * rx;
        xb1 += G * qy + rq * ry;
        xb2 += G * qz + rq * rz;
        if (visc) {
            const double mubar = pibar_tot * (-alpha * cm + 2.0 * beta * mu) / rm;
            cbar_a += 0.5 * pibar_tot * (-alpha * mu / rm);
            rhobar_pi += 0.5 * pibar_tot * (-Pi / rm);
            albar += pibar_own * (-cm * mu / rm);
            const double dotbar = mubar * ha / den;
            const double r2bar = -mubar * ha * dot / (den * den);
            xb0 += dotbar * vx + 2.0 * r2bar * rx;
            xb1 += dotbar * vy + 2.0 * r2bar * ry;
            xb2 += dotbar * vz + 2.0 * r2bar * rz;
            ub0 += dotbar * rx;
            ub1 += dotbar * ry;
            ub2 += dotbar * rz;
        }
    }
    if (bad_h) atomicOr(status, 32);
    // P = p * (1/rho^2); p = (pow(rho/ro, gamma) - 1) B; cs = pow(rho/ro, (gamma-1)/2) co
    const double Bc = co * co * ro / gamma;
    const double ratio = rhoa / ro;
    const double dp = Bc * gamma * pow(ratio, gamma - 1.0) / ro;
    const double g2 = 0.5 * (gamma - 1.0);
    const double dc = co * g2 * pow(ratio, g2 - 1.0) / ro;
    const double inv2 = 1.0 / (rhoa * rhoa);
    rhobar[t] = rhobar_pi + Pbar * (-2.0 * pa * inv2 / rhoa) + Pbar * inv2 * dp + cbar_a * dc;
    xbar[3 * t] = xb0;
    xbar[3 * t + 1] = xb1;
    xbar[3 * t + 2] = xb2;
    ubar[3 * t] = ub0;
    ubar[3 * t + 1] = ub1;
    ubar[3 * t + 2] = ub2;
    alphabar[t] = albar;
}

// density part: xbar_a += sum_b (rhobar_a m_b + rhobar_b m_a) gradW_ab, then
// everything back to original order
__global__ void __launch_bounds__(BT_THREADS)
    k_vjp_density(int64_t n, const int32_t *__restrict__ perm, const int32_t *__restrict__ counts,
                  const int64_t *__restrict__ slice_off, const int32_t *__restrict__ indices,
                  const double *__restrict__ rec, const double *__restrict__ rhobar, int dim,
                  const double *__restrict__ xbar_s, const double *__restrict__ ubar_s,
                  double *__restrict__ xbar, double *__restrict__ ybar, double *__restrict__ zbar,
                  double *__restrict__ ubar, double *__restrict__ vbar, double *__restrict__ wbar)
{
    const int64_t t = (int64_t)sph_xcd_tile(blockIdx.x, gridDim.x) * BT_THREADS + threadIdx.x;
    if (t >= n) return;
    const double *ra = rec + t * SPH_REC;
    const double xa = ra[SPH_R_X], ya = ra[SPH_R_Y], za = ra[SPH_R_Z], ma = ra[SPH_R_M];
    const double ha = ra[SPH_R_H];
    const double fac = sph_h_dim(ha, dim) * sph_cubic_fac1(dim);
    const double rba = rhobar[t];
    double xb0 = xbar_s[3 * t], xb1 = xbar_s[3 * t + 1], xb2 = xbar_s[3 * t + 2];
    const int32_t *row = indices + slice_off[t >> 6] + (t & 63);
    const int32_t e = counts[t];
    for (int32_t k = 0; k < e; ++k) {
        const int32_t b = row[(int64_t)k << 6];
        if (b == (int32_t)t) continue;
        const double *rb = rec + (int64_t)b * SPH_REC;
        const double rx = xa - rb[SPH_R_X], ry = ya - rb[SPH_R_Y], rz = za - rb[SPH_R_Z];
        const double rab = sqrt(rx * rx + ry * ry + rz * rz);
        double G, dG;
        cubic_g_dg(rab, ha, fac, G, dG);
        const double c = (rba * rb[SPH_R_M] + rhobar[b] * ma) * G;
        xb0 += c * rx;
        xb1 += c * ry;
        xb2 += c * rz;
    }
    const int32_t i = perm[t];
    xbar[i] = xb0;
    ybar[i] = xb1;
    if (zbar) zbar[i] = xb2;
    ubar[i] = ubar_s[3 * t];
    vbar[i] = ubar_s[3 * t + 1];
    if (wbar) wbar[i] = ubar_s[3 * t + 2];
}

// deterministic sum of n doubles by one block
__global__ void __launch_bounds__(1024)
    k_sum_f64(const double *__restrict__ v, int64_t n, double *__restrict__ out)
{
    __shared__ double part[1024 / 64];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 1024) s += v[i];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int w = 0; w < 1024 / 64; ++w) tot += part[w];
        *out = tot;
    }
}

// ---------------------------------------------------------------------------
// workspace
// ---------------------------------------------------------------------------
namespace {

struct BatchWork {
    GridDev *grids;
    int64_t *ncell;
    int32_t *cell_start, *cell_of, *rank, *perm, *cells, *slot, *tmp;
    // simulation of each cell key (written by the binning for the keys it
    // fills): the list count pass finds a cell's grid without going through
    // its first particle
    int32_t *key_sim;
    double *rec;
    int32_t *counts;
    int64_t *slice_off;
    int32_t *indices;
    // reverse mode, sorted order: cotangent of the accelerations, adjoints
    // of x and u (3 each), of rho, and the per-particle alpha terms
    double *abar, *xbar, *ubar, *rhobar, *alphabar;
    // arrival counter of the persistent evaluation's grid barrier
    int32_t *arrived;
    size_t bytes;
};

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

// carve the caller's workspace (base may be NULL to only compute the size)
BatchWork carve(const sph_batch_desc *d, char *base)
{
    BatchWork w;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        char *p = base ? base + o : nullptr;
        o += al256(bytes);
        return p;
    };
    const int64_t n = d->n, kc = d->key_cap, nsl = (n + 63) / 64;
    w.grids = (GridDev *)take(sizeof(GridDev) * (size_t)d->nsims);
    w.ncell = (int64_t *)take(sizeof(int64_t) * (size_t)d->nsims);
    w.cell_start = (int32_t *)take(sizeof(int32_t) * (size_t)(kc + 1));
    w.cell_of = (int32_t *)take(sizeof(int32_t) * (size_t)n);
    w.rank = (int32_t *)take(sizeof(int32_t) * (size_t)n);
    w.perm = (int32_t *)take(sizeof(int32_t) * (size_t)n);
    w.cells = (int32_t *)take(sizeof(int32_t) * (size_t)kc);
    w.slot = (int32_t *)take(sizeof(int32_t) * (size_t)(kc + 1));
    w.tmp = (int32_t *)take(sizeof(int32_t) * (size_t)n);
    w.rec = (double *)take(sizeof(double) * SPH_REC * (size_t)n);  // float records when f32
    w.counts = (int32_t *)take(sizeof(int32_t) * (size_t)(n + 1));
    w.slice_off = (int64_t *)take(sizeof(int64_t) * 2 * (size_t)(nsl + 1));
    w.indices = (int32_t *)take(sizeof(int32_t) * (size_t)d->list_cap);
    w.abar = (double *)take(sizeof(double) * 3 * (size_t)n);
    w.xbar = (double *)take(sizeof(double) * 3 * (size_t)n);
    w.ubar = (double *)take(sizeof(double) * 3 * (size_t)n);
    w.rhobar = (double *)take(sizeof(double) * (size_t)n);
    w.alphabar = (double *)take(sizeof(double) * (size_t)n);
    w.arrived = (int32_t *)take(sizeof(int32_t) * 64);
    w.key_sim = (int32_t *)take(sizeof(int32_t) * (size_t)kc);
    w.bytes = o;
    return w;
}

int check_desc(sph_ctx *ctx, const sph_batch_desc *d)
{
    SPH_REQUIRE(ctx, d, "batch: null descriptor");
    SPH_REQUIRE(ctx, d->nsims >= 1 && (d->dim == 2 || d->dim == 3), "batch: bad nsims/dim");
    SPH_REQUIRE(ctx, d->n >= 1 && d->n < ((int64_t)1 << 31), "batch: bad n");
    SPH_REQUIRE(ctx, d->key_cap >= 2 && d->key_cap < ((int64_t)1 << 31) - 1 && d->list_cap >= 1 &&
                         d->max_cells >= 1 && d->max_cells <= d->key_cap,
                "batch: bad capacities");
    SPH_REQUIRE(ctx, d->maxch >= 1 && d->maxch <= 16, "batch: bad maxch");
    SPH_REQUIRE(ctx, d->f32 == 0 || d->f32 == 1, "batch: bad f32 flag");
    SPH_REQUIRE(ctx, d->uniform_h == 0 || d->uniform_h == 1, "batch: bad uniform_h flag");
    SPH_REQUIRE(ctx, d->op_threads == 0 || d->op_threads == 64 || d->op_threads == 128 ||
                         d->op_threads == 256,
                "batch: op_threads must be 64, 128 or 256 (or 0)");
    SPH_REQUIRE(ctx, d->op_blocks_per_cu >= 0 && d->op_blocks_per_cu <= 64,
                "batch: op_blocks_per_cu must be 0..64");
    SPH_REQUIRE(ctx, d->nbr_tile == 0 || d->nbr_tile == 1 || d->nbr_tile == 2 || d->nbr_tile == 4,
                "batch: nbr_tile must be 1, 2 or 4 (or 0)");
    SPH_REQUIRE(ctx, d->cells_per_block >= 0 && d->cells_per_block <= 64,
                "batch: cells_per_block must be 0..64");
    SPH_REQUIRE(ctx, d->persistent == 0 || d->persistent == 1, "batch: persistent must be 0 or 1");
    SPH_REQUIRE(ctx, d->staged == 0 || d->staged == 1, "batch: staged must be 0 or 1");
    SPH_REQUIRE(ctx, d->sim_off && d->sim_of && d->x && d->y && d->z && d->u && d->v && d->w &&
                         d->h && d->m && d->rho && d->p && d->cs && d->x0 && d->y0 && d->u0 &&
                         d->v0 && d->ax && d->ay && d->au && d->av && d->dt_fac && d->status &&
                         d->pairs,
                "batch: null array");
    SPH_REQUIRE(ctx, d->dim == 2 || (d->z0 && d->w0 && d->az && d->aw),
                "batch: 3D needs z0, w0, az, aw");
    SPH_REQUIRE(ctx, d->work, "batch: null workspace");
    return SPH_OK;
}

// binning, sorted records and neighbour lists at the current positions
int batch_lists(sph_ctx *ctx, const sph_batch_desc *d, const BatchWork &w)
{
    const int64_t n = d->n;
    int rc;
    // -- binning into the per-simulation grids
    hipLaunchKernelGGL(k_batch_grid, dim3(d->nsims), dim3(GRID_THREADS), 0, ctx->stream,
                       d->sim_off, d->x, d->y, d->z, d->h, d->radius_scale, d->uniform_h,
                       w.grids, w.ncell, d->status);
    SPH_LAUNCH_CHECK(ctx, "k_batch_grid");
    hipLaunchKernelGGL(k_batch_key_base, dim3(1), dim3(1024), 0, ctx->stream, d->nsims, w.ncell,
                       d->key_cap, w.grids, d->status);
    SPH_LAUNCH_CHECK(ctx, "k_batch_key_base");
    // cell counts into w.slot (sph_bin_sort turns them into w.cell_start)
    hipLaunchKernelGGL(k_fill_i32, dim3(sph_blocks(d->key_cap + 1, BT_THREADS)),
                       dim3(BT_THREADS), 0, ctx->stream, w.slot, d->key_cap + 1, 0);
    SPH_LAUNCH_CHECK(ctx, "k_fill_i32");
    hipLaunchKernelGGL(k_batch_keys, dim3(sph_blocks(n, BT_THREADS)), dim3(BT_THREADS), 0,
                       ctx->stream, w.grids, d->sim_of, d->x, d->y, d->z, n, d->key_cap,
                       w.slot, w.cell_of, w.rank, w.key_sim, d->status);
    SPH_LAUNCH_CHECK(ctx, "k_batch_keys");
    const int32_t trash = (int32_t)(d->key_cap - 1);
    rc = sph_bin_sort(ctx, n, d->key_cap, w.cell_of, w.rank, w.cell_start, w.perm, w.cells,
                      w.slot, w.tmp, trash, 0);
    if (rc) return rc;
    const int32_t *ncell_dev = w.slot + d->key_cap;
    hipLaunchKernelGGL(k_batch_check_cells, dim3(1), dim3(64), 0, ctx->stream, ncell_dev,
                       d->max_cells, d->status);
    SPH_LAUNCH_CHECK(ctx, "k_batch_check_cells");
    // -- sorted records of the fields the operators read (written from the
    // original order into row rank[i]: coalesced reads)
    // in 2D the z and w fields of the records stay zero (the workspace
    // record area is cleared by sph_batch_init_workspace)
    const double *src[8] = {d->x, d->y, d->m, d->u, d->v, d->h, d->z, d->w};
    const int fld[8] = {SPH_R_X, SPH_R_Y, SPH_R_M, SPH_R_U, SPH_R_V, SPH_R_H, SPH_R_Z, SPH_R_W};
    const int nf = d->dim == 3 ? 8 : 6;
    float *rec32 = d->f32 ? (float *)w.rec : nullptr;
    if (rec32) {
        sph_prop_list_f32 pl;
        memset(&pl, 0, sizeof(pl));
        pl.dst_stride = SPH_REC;
        pl.nprops = nf;
        for (int k = 0; k < nf; ++k) {
            pl.src[k] = src[k];
            pl.dst[k] = rec32 + fld[k];
        }
        rc = sph_scatter_sorted_f32(ctx, n, w.rank, &pl);
    } else {
        sph_prop_list pl;
        memset(&pl, 0, sizeof(pl));
        pl.dst_stride = SPH_REC;
        pl.nprops = nf;
        for (int k = 0; k < nf; ++k) {
            pl.src[k] = src[k];
            pl.dst[k] = w.rec + fld[k];
        }
        rc = sph_scatter_sorted(ctx, n, w.rank, &pl);
    }
    if (rc) return rc;
    // -- neighbour lists
    rc = sph_nnps_build_batch(ctx, w.grids, d->sim_of, w.key_sim, rec32 ? nullptr : w.rec, rec32, d->x,
                              d->y, d->z, w.perm, w.cell_start, w.cells, ncell_dev,
                              d->max_cells, n, d->radius_scale, d->maxch, w.counts, w.slice_off,
                              w.indices, d->list_cap, d->status, d->pairs, trash,
                              d->cells_per_block > 0 ? d->cells_per_block : 1);
    return rc;
}

// one evaluation of the operation list at the current positions, kernel by
// kernel (the reverse mode drives these pieces)
int batch_eval(sph_ctx *ctx, const sph_batch_desc *d, const BatchWork &w)
{
    const int64_t n = d->n;
    int rc = batch_lists(ctx, d, w);
    if (rc) return rc;
    float *rec32 = d->f32 ? (float *)w.rec : nullptr;
    // -- SPHRho, Tait (non-integrating: outputs start from 0), MomentumEquation
    // (dt_fac reset per calc), PositionStepping
    double *az = d->dim == 3 ? d->aw : nullptr;
    if (rec32) {
        rc = sph_summation_density_f32(ctx, d->dim, n, n, w.perm, w.counts, w.slice_off,
                                       w.indices, rec32, rec32, d->uniform_h, 0, d->rho,
                                       rec32 + SPH_R_RHO, SPH_REC);
        if (rc) return rc;
        rc = sph_tait_records_f32(ctx, n, w.perm, rec32, d->ro, d->co, d->gamma, d->p, d->cs);
        if (rc) return rc;
        rc = sph_fill(ctx, n, 0.0, d->dt_fac);
        if (rc) return rc;
        // ns = 0: P/rho^2 is already in the records
        rc = sph_momentum_av_f32(ctx, d->dim, n, n, w.perm, w.counts, w.slice_off, w.indices,
                                 rec32, rec32, 0, d->uniform_h, d->alpha, d->beta, d->eta, 0,
                                 d->au, d->av, az, d->dt_fac);
    } else {
        rc = sph_summation_density(ctx, d->dim, n, n, w.perm, w.counts, w.slice_off, w.indices,
                                   w.rec, w.rec, d->uniform_h, 0, d->rho, w.rec + SPH_R_RHO,
                                   SPH_REC);
        if (rc) return rc;
        rc = sph_tait_records(ctx, n, w.perm, w.rec, d->ro, d->co, d->gamma, d->p, d->cs);
        if (rc) return rc;
        rc = sph_fill(ctx, n, 0.0, d->dt_fac);
        if (rc) return rc;
        rc = sph_momentum_av(ctx, d->dim, n, n, w.perm, w.counts, w.slice_off, w.indices, w.rec,
                             w.rec, 0, d->uniform_h, d->alpha, d->beta, d->eta, 0, d->au, d->av,
                             az, d->dt_fac);
    }
    if (rc) return rc;
    return sph_position_stepping(ctx, n, n, d->u, d->v, d->dim == 3 ? d->w : nullptr, d->ax,
                                 d->ay, d->dim == 3 ? d->az : nullptr);
}

int batch_stage(sph_ctx *ctx, const sph_batch_desc *d, double dt)
{
    sph_step_props sp;
    memset(&sp, 0, sizeof(sp));
    double *x[6] = {d->u, d->v, d->x, d->y, d->w, d->z};
    const double *x0[6] = {d->u0, d->v0, d->x0, d->y0, d->w0, d->z0};
    const double *a[6] = {d->au, d->av, d->ax, d->ay, d->aw, d->az};
    sp.nprops = d->dim == 3 ? 6 : 4;
    for (int k = 0; k < sp.nprops; ++k) {
        sp.x[k] = x[k];
        sp.x0[k] = x0[k];
        sp.a[k] = a[k];
    }
    if (!d->dt_sim) return sph_stage_step(ctx, d->n, &sp, dt);
    hipLaunchKernelGGL(k_batch_stage, dim3(sph_blocks(d->n, BT_THREADS)), dim3(BT_THREADS), 0,
                       ctx->stream, d->n, d->sim_of, d->dt_sim, dt, sp);
    SPH_LAUNCH_CHECK(ctx, "k_batch_stage");
    return SPH_OK;
}

// blocks of k_batch_fused_eval the device holds at once at this block size
template <bool UH, typename T, int NT>
static int64_t fused_resident_blocks(sph_ctx *ctx, unsigned threads)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_fused_eval<UH, T, NT>,
                                                     (int)threads, 0) != hipSuccess)
        return 0;
    return (int64_t)per_cu * sph_cu_count(ctx);
}

template <typename T, int NT>
void launch_fused(sph_ctx *ctx, const sph_batch_desc *d, const BatchWork &w, unsigned grid,
                  unsigned threads, int64_t ntiles, double scale, int save)
{
    const int64_t n = d->n;
    T *rec = (T *)w.rec;
    const T B = T(d->co) * T(d->co) * T(d->ro) / T(d->gamma);  // eos_funcs.pyx:90
    FusedState st;
    double *x[3] = {d->x, d->y, d->z}, *u[3] = {d->u, d->v, d->w};
    double *x0[3] = {d->x0, d->y0, d->z0}, *u0[3] = {d->u0, d->v0, d->w0};
    double *ax[3] = {d->ax, d->ay, d->az}, *au[3] = {d->au, d->av, d->aw};
    for (int c = 0; c < 3; ++c) {
        st.x[c] = x[c];
        st.u[c] = u[c];
        st.x0[c] = x0[c];
        st.u0[c] = u0[c];
        st.ax[c] = ax[c];
        st.au[c] = au[c];
    }
    st.dt_fac = d->dt_fac;
    if (d->staged) {
        // records staged in LDS: one wavefront per 64-row slice, two launches
        const unsigned ns = (unsigned)((n + STG_THREADS - 1) / STG_THREADS);
        const double *ox = d->f32 ? d->x : nullptr, *oy = d->f32 ? d->y : nullptr,
                     *oz = d->f32 ? d->z : nullptr;
        auto dks = d->uniform_h ? k_batch_density_eos_staged<true, T>
                                : k_batch_density_eos_staged<false, T>;
        hipLaunchKernelGGL(dks, dim3(ns), dim3(STG_THREADS), 0, ctx->stream, d->dim, n, w.grids,
                           d->sim_of, w.perm, w.cell_start, w.counts, w.slice_off, w.indices,
                           rec, ox, oy, oz, d->radius_scale, T(d->ro), T(d->co), T(d->gamma), B,
                           d->rho, d->p, d->cs);
        auto mks = d->uniform_h ? k_batch_momentum_stage_staged<true, T>
                                : k_batch_momentum_stage_staged<false, T>;
        hipLaunchKernelGGL(mks, dim3(ns), dim3(STG_THREADS), 0, ctx->stream, d->dim, n, w.grids,
                           d->sim_of, w.perm, w.cell_start, w.counts, w.slice_off, w.indices,
                           (const T *)rec, ox, oy, oz, d->radius_scale, T(d->alpha),
                           T(d->beta), T(d->eta), st, d->dt_sim, scale, save);
        return;
    }
    // the whole evaluation in one launch of at most the resident blocks (two
    // launches when the occupancy query fails)
    const int64_t cap = !d->persistent ? 0
                        : d->uniform_h ? fused_resident_blocks<true, T, NT>(ctx, threads)
                                       : fused_resident_blocks<false, T, NT>(ctx, threads);
    if (cap > 0) {
        const int64_t g = (int64_t)grid < cap ? (int64_t)grid : cap;
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, ctx->stream, w.arrived, 1, 0);
        auto fk = d->uniform_h ? k_batch_fused_eval<true, T, NT> : k_batch_fused_eval<false, T, NT>;
        hipLaunchKernelGGL(fk, dim3((unsigned)g), dim3(threads), 0, ctx->stream, d->dim, n,
                           ntiles, w.perm, w.counts, w.slice_off, w.indices, rec, T(d->ro),
                           T(d->co), T(d->gamma), B, d->rho, d->p, d->cs, T(d->alpha), T(d->beta),
                           T(d->eta), st, d->sim_of, d->dt_sim, scale, save, w.arrived, d->status,
                           FUSED_SPIN_LIMIT);
        return;
    }
    auto dk = d->uniform_h ? k_batch_density_eos<true, T, 2 * NT>
                           : k_batch_density_eos<false, T, 2 * NT>;
    hipLaunchKernelGGL(dk, dim3(grid), dim3(threads), 0, ctx->stream, d->dim, n, ntiles, w.perm,
                       w.counts, w.slice_off, w.indices, rec, T(d->ro), T(d->co), T(d->gamma), B,
                       d->rho, d->p, d->cs);
    auto mk = d->uniform_h ? k_batch_momentum_stage<true, T, NT>
                           : k_batch_momentum_stage<false, T, NT>;
    hipLaunchKernelGGL(mk, dim3(grid), dim3(threads), 0, ctx->stream, d->dim, n, ntiles, w.perm,
                       w.counts, w.slice_off, w.indices, (const T *)rec, T(d->alpha),
                       T(d->beta), T(d->eta), st, d->sim_of, d->dt_sim, scale, save);
}

// one evaluation and its stage through the fused kernels: x <- x0 + a*(scale
// dt); save: x0 <- x first (save_initial_arrays, solver/integrator.py:262-285)
int batch_eval_stage_fused(sph_ctx *ctx, const sph_batch_desc *d, const BatchWork &w,
                           double scale, int save)
{
    int rc = batch_lists(ctx, d, w);
    if (rc) return rc;
    const unsigned threads = d->op_threads > 0 ? (unsigned)d->op_threads : FUSED_MAX_THREADS;
    const int64_t ntiles = (d->n + threads - 1) / threads;
    int64_t grid = ntiles;
    if (d->op_blocks_per_cu > 0) {
        const int64_t g = (int64_t)d->op_blocks_per_cu * sph_cu_count(ctx);
        if (g < grid) grid = g;
    }
    const int nt = d->nbr_tile > 0 ? d->nbr_tile : 2;
    if (d->f32) {
        if (nt == 1) launch_fused<float, 1>(ctx, d, w, (unsigned)grid, threads, ntiles, scale, save);
        else if (nt == 2) launch_fused<float, 2>(ctx, d, w, (unsigned)grid, threads, ntiles, scale, save);
        else launch_fused<float, 4>(ctx, d, w, (unsigned)grid, threads, ntiles, scale, save);
    } else {
        if (nt == 1) launch_fused<double, 1>(ctx, d, w, (unsigned)grid, threads, ntiles, scale, save);
        else if (nt == 2) launch_fused<double, 2>(ctx, d, w, (unsigned)grid, threads, ntiles, scale, save);
        else launch_fused<double, 4>(ctx, d, w, (unsigned)grid, threads, ntiles, scale, save);
    }
    SPH_LAUNCH_CHECK(ctx, "k_batch_density_eos / k_batch_momentum_stage");
    return SPH_OK;
}

int batch_rk2(sph_ctx *ctx, const sph_batch_desc *d, const BatchWork &w, double dt)
{
    // RK2 midpoint: X1 = X0 + (dt/2) F(X0) (saving X0 first), X = X0 + dt F(X1)
    int rc = batch_eval_stage_fused(ctx, d, w, dt / 2, 1);
    if (rc) return rc;
    return batch_eval_stage_fused(ctx, d, w, dt, 0);
}

struct BatchGraph {
    hipGraph_t graph;
    hipGraphExec_t exec;
};

}  // namespace

extern "C" int sph_batch_workspace_size(sph_ctx *ctx, const sph_batch_desc *d, int64_t *bytes)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, d && bytes, "sph_batch_workspace_size: null argument");
    *bytes = (int64_t)carve(d, nullptr).bytes;
    return SPH_OK;
}

extern "C" int sph_batch_rk2_step(sph_ctx *ctx, const sph_batch_desc *d, double dt)
{
    if (!ctx) return SPH_EINVAL;
    int rc = check_desc(ctx, d);
    if (rc) return rc;
    BatchWork w = carve(d, (char *)d->work);
    SPH_REQUIRE(ctx, d->work_bytes >= (int64_t)w.bytes, "batch: workspace of %lld bytes < %zu",
                (long long)d->work_bytes, w.bytes);
    return batch_rk2(ctx, d, w, dt);
}

static int batch_prep(sph_ctx *ctx, const sph_batch_desc *d, BatchWork *w)
{
    if (!ctx) return SPH_EINVAL;
    int rc = check_desc(ctx, d);
    if (rc) return rc;
    *w = carve(d, (char *)d->work);
    SPH_REQUIRE(ctx, d->work_bytes >= (int64_t)w->bytes, "batch: workspace of %lld bytes < %zu",
                (long long)d->work_bytes, w->bytes);
    return SPH_OK;
}

extern "C" int sph_batch_init_workspace(sph_ctx *ctx, const sph_batch_desc *d)
{
    BatchWork w;
    int rc = batch_prep(ctx, d, &w);
    if (rc) return rc;
    // zero the records (2D reads z = w = 0 from them) with a fill kernel
    const int64_t nwords = (int64_t)(SPH_REC * (size_t)d->n * sizeof(double) / sizeof(int32_t));
    hipLaunchKernelGGL(k_fill_i32, dim3(sph_blocks(nwords, BT_THREADS)), dim3(BT_THREADS), 0,
                       ctx->stream, (int32_t *)w.rec, nwords, 0);
    SPH_LAUNCH_CHECK(ctx, "k_fill_i32");
    return SPH_OK;
}

extern "C" int sph_batch_save_initial(sph_ctx *ctx, const sph_batch_desc *d)
{
    BatchWork w;
    int rc = batch_prep(ctx, d, &w);
    if (rc) return rc;
    double *dst[6] = {d->u0, d->v0, d->x0, d->y0, d->w0, d->z0};
    const double *src[6] = {d->u, d->v, d->x, d->y, d->w, d->z};
    for (int k = 0; k < (d->dim == 3 ? 6 : 4); ++k) {
        rc = sph_copy(ctx, d->n, src[k], dst[k]);
        if (rc) return rc;
    }
    return SPH_OK;
}

extern "C" int sph_batch_eval(sph_ctx *ctx, const sph_batch_desc *d)
{
    BatchWork w;
    int rc = batch_prep(ctx, d, &w);
    if (rc) return rc;
    return batch_eval(ctx, d, w);
}

extern "C" int sph_batch_stage(sph_ctx *ctx, const sph_batch_desc *d, double dt)
{
    BatchWork w;
    int rc = batch_prep(ctx, d, &w);
    if (rc) return rc;
    return batch_stage(ctx, d, dt);
}

extern "C" int sph_batch_vjp(sph_ctx *ctx, const sph_batch_desc *d, const double *abar_u,
                             const double *abar_v, const double *abar_w, double *xbar,
                             double *ybar, double *zbar, double *ubar, double *vbar,
                             double *wbar, double *alpha_bar)
{
    BatchWork w;
    int rc = batch_prep(ctx, d, &w);
    if (rc) return rc;
    SPH_REQUIRE(ctx, abar_u && abar_v && xbar && ybar && ubar && vbar && alpha_bar,
                "sph_batch_vjp: null argument");
    SPH_REQUIRE(ctx, !d->f32, "sph_batch_vjp: the reverse mode runs in FP64 only");
    SPH_REQUIRE(ctx, d->dim == 2 || (abar_w && zbar && wbar), "sph_batch_vjp: 3D needs w/z");
    const int64_t n = d->n;
    hipLaunchKernelGGL(k_gather3, dim3(sph_blocks(n, BT_THREADS)), dim3(BT_THREADS), 0,
                       ctx->stream, n, (const int32_t *)w.perm, abar_u, abar_v,
                       d->dim == 3 ? abar_w : nullptr, w.abar);
    SPH_LAUNCH_CHECK(ctx, "k_gather3");
    hipLaunchKernelGGL(k_vjp_momentum, dim3(sph_blocks(n, BT_THREADS)), dim3(BT_THREADS), 0,
                       ctx->stream, n, (const int32_t *)w.counts, (const int64_t *)w.slice_off,
                       (const int32_t *)w.indices, (const double *)w.rec,
                       (const double *)w.abar, d->dim, d->alpha, d->beta, d->eta, d->ro, d->co,
                       d->gamma, w.xbar, w.ubar, w.rhobar, w.alphabar, d->status);
    SPH_LAUNCH_CHECK(ctx, "k_vjp_momentum");
    hipLaunchKernelGGL(k_vjp_density, dim3(sph_blocks(n, BT_THREADS)), dim3(BT_THREADS), 0,
                       ctx->stream, n, (const int32_t *)w.perm, (const int32_t *)w.counts,
                       (const int64_t *)w.slice_off, (const int32_t *)w.indices,
                       (const double *)w.rec, (const double *)w.rhobar, d->dim,
                       (const double *)w.xbar, (const double *)w.ubar, xbar, ybar,
                       d->dim == 3 ? zbar : nullptr, ubar, vbar, d->dim == 3 ? wbar : nullptr);
    SPH_LAUNCH_CHECK(ctx, "k_vjp_density");
    hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(1024), 0, ctx->stream, (const double *)w.alphabar,
                       n, alpha_bar);
    SPH_LAUNCH_CHECK(ctx, "k_sum_f64");
    return SPH_OK;
}

extern "C" int sph_batch_graph_create(sph_ctx *ctx, const sph_batch_desc *d, double dt,
                                      int nsteps, void **graph_out)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, graph_out && nsteps >= 1, "sph_batch_graph_create: bad argument");
    *graph_out = nullptr;
    int rc = check_desc(ctx, d);
    if (rc) return rc;
    SPH_REQUIRE(ctx, ctx->stream != nullptr, "sph_batch_graph_create: the context needs a stream");
    BatchWork w = carve(d, (char *)d->work);
    SPH_REQUIRE(ctx, d->work_bytes >= (int64_t)w.bytes, "batch: workspace of %lld bytes < %zu",
                (long long)d->work_bytes, w.bytes);
    // the scratch must not move while a graph refers to it: the caller runs
    // one eager step first (which sizes it) and the context stays frozen
    ctx->frozen = 1;
    BatchGraph *g = (BatchGraph *)calloc(1, sizeof(BatchGraph));
    if (!g) return sph_set_error(ctx, SPH_ENOMEM, "graph handle");
    SPH_HIP_CHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    for (int s = 0; s < nsteps && rc == SPH_OK; ++s) rc = batch_rk2(ctx, d, w, dt);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (rc != SPH_OK || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        free(g);
        if (rc != SPH_OK) return rc;
        return sph_set_error(ctx, SPH_EHIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        free(g);
        return sph_set_error(ctx, SPH_EHIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    }
    g->graph = graph;
    *graph_out = g;
    return SPH_OK;
}

extern "C" int sph_batch_graph_launch(sph_ctx *ctx, void *graph)
{
    if (!ctx) return SPH_EINVAL;
    SPH_REQUIRE(ctx, graph, "sph_batch_graph_launch: null graph");
    SPH_HIP_CHECK(ctx, hipGraphLaunch(((BatchGraph *)graph)->exec, ctx->stream));
    return SPH_OK;
}

extern "C" int sph_batch_graph_destroy(sph_ctx *ctx, void *graph)
{
    if (!ctx) return SPH_EINVAL;
    if (!graph) return SPH_OK;
    BatchGraph *g = (BatchGraph *)graph;
    SPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    free(g);
    return SPH_OK;
}

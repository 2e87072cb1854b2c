// Internal definitions shared by the HIP translation units of libsphhip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

#include "sphhip.h"

#define SPH_VERSION_STRING "pysph_amd libsphhip 0.1 (gfx950)"

struct sph_ctx {
    int device;
    hipStream_t stream;
    char err[512];
    // growable device scratch, carved by the calls that need it
    void *scratch;
    size_t scratch_bytes;
    // pinned host staging for small readbacks
    void *pinned;
    size_t pinned_bytes;
    // per-axis cell-key tables of the current grid (device, int32)
    int32_t *key_tab;
    int64_t key_tab_len;
    int32_t key_dims[3], key_bits[3], key_order;
    // set once a graph refers to the scratch: it may no longer be reallocated
    int frozen;
    // compute units of the device (queried on first use)
    int ncu;
    // device address of the file-scope partials array of sph_extrema_noalloc
    double *extrema_scratch;
};

int sph_cu_count(sph_ctx *ctx);
int sph_extrema_scratch_setup(sph_ctx *ctx);

int sph_set_error(sph_ctx *ctx, int code, const char *fmt, ...);
int sph_ensure_scratch(sph_ctx *ctx, size_t bytes);
int sph_ensure_pinned(sph_ctx *ctx, size_t bytes);
int sph_upload_key_tables(sph_ctx *ctx, const sph_grid *g);

#define SPH_HIP_CHECK(ctx, expr)                                                              \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return sph_set_error((ctx), SPH_EHIP, "%s failed: %s", #expr,                    \
                                 hipGetErrorString(e_));                                      \
    } while (0)

#define SPH_LAUNCH_CHECK(ctx, what)                                                           \
    do {                                                                                      \
        hipError_t e_ = hipGetLastError();                                                    \
        if (e_ != hipSuccess)                                                                 \
            return sph_set_error((ctx), SPH_EHIP, "launch of %s failed: %s", (what),         \
                                 hipGetErrorString(e_));                                      \
    } while (0)

#define SPH_REQUIRE(ctx, cond, ...)                                                           \
    do {                                                                                      \
        if (!(cond)) return sph_set_error((ctx), SPH_EINVAL, __VA_ARGS__);                    \
    } while (0)

static inline unsigned sph_blocks(int64_t n, int threads)
{
    int64_t b = (n + threads - 1) / threads;
    return (unsigned)(b > 0 ? b : 1);
}

// Blocks are dispatched round-robin over the 8 XCDs of an MI355X, each XCD
// with its own L2.  Remap block b so that every XCD walks a contiguous range
// of tiles (of the cell-sorted order): neighbouring particles then share
// one L2.  Bijective on [0, nb).
#define SPH_NXCD 8
__device__ __forceinline__ unsigned sph_xcd_tile(unsigned b, unsigned nb)
{
    const unsigned q = nb / SPH_NXCD, r = nb % SPH_NXCD;
    const unsigned x = b % SPH_NXCD, k = b / SPH_NXCD;
    return x * q + (x < r ? x : r) + k;
}

// A cell grid as the kernels see it.  Cell keys come from the per-axis
// tables tab = tx | ty | tz (Morton or linear order) or, with tab == NULL,
// are key_base + the linear index (the per-simulation grids of a batch).
struct GridDev {
    double cs;
    int cmin0, cmin1, cmin2;
    int d0, d1, d2;
    const int32_t *tab;
    int32_t key_base;
};

__device__ __forceinline__ int cell_coord(double v, double cs) { return (int)floor(v / cs); }

__device__ __forceinline__ int32_t cell_key(const GridDev &g, int cx, int cy, int cz)
{
    if (g.tab) return g.tab[cx] + g.tab[g.d0 + cy] + g.tab[g.d0 + g.d1 + cz];
    return g.key_base + (cx * g.d1 + cy) * g.d2 + cz;
}

__global__ void k_fill_i32(int32_t *__restrict__ p, int64_t n, int32_t v);

// Counting-sort tail shared by the single-grid and batched binning: given
// cell_of[n] and, per particle, its slot handed out by the atomic count
// (in rank), and the per-cell counts in cell_start[0..nc), produce
// cell_start (exclusive scan), perm/rank (ascending index inside every
// cell), slot[0..nc] (exclusive scan of the non-empty flags; slot[nc] = the
// number of non-empty cells) and cells[] (their keys).  tmp holds n ints.
// The members of cell `trash` (-1: none) are left unordered.  Scan partials
// go to the context scratch from scratch_off on.  ncells_copy (device, may be
// NULL) receives the number of non-empty cells too.
int sph_bin_sort(sph_ctx *ctx, int64_t n, int64_t nc, const int32_t *cell_of, int32_t *rank,
                 int32_t *cell_start, int32_t *perm, int32_t *cells, int32_t *slot, int32_t *tmp,
                 int32_t trash, size_t scratch_off, int32_t *ncells_copy = nullptr);
// end of the context scratch sph_bin_sort uses from scratch_off on (scan
// partials and the list of cells with many members); a caller that keeps
// pointers into the scratch below scratch_off ensures this size first
size_t sph_bin_sort_scratch_end(int64_t n, int64_t nc, size_t scratch_off);

// Neighbour lists of a batch of simulations packed in one array, without
// host synchronisation (capturable in a graph): grids[sim_of[i]] is the grid
// of original particle i, *ncell_dev the number of non-empty cells (at most
// max_cells).  Lists needing more than cap slots set bit 0 of *status and
// are not written; nnz is added to *nnz_acc.  Particles in cell `trash` get
// empty rows.  With rec32 != NULL the records are float (rec unused) and
// ox, oy, oz are the double coordinates in original order.
int sph_nnps_build_batch(sph_ctx *ctx, const GridDev *grids, const int32_t *sim_of,
                         const int32_t *key_sim,
                         const double *rec, const float *rec32, const double *ox,
                         const double *oy, const double *oz, const int32_t *perm,
                         const int32_t *cell_start,
                         const int32_t *cells, const int32_t *ncell_dev, int64_t max_cells,
                         int64_t n, double radius_scale, int maxch, int32_t *counts,
                         int64_t *slice_off, int32_t *indices, int64_t cap, int32_t *status,
                         int64_t *nnz_acc, int32_t trash, int cells_per_block);

// TaitEquation over sorted records (all particles real, no accumulation):
// p, cs into the original-order arrays and the record fields, P/rho^2 into
// SPH_R_PRHO2 (so a following sph_momentum_av can skip it with ns = 0).
int sph_tait_records(sph_ctx *ctx, int64_t n, const int32_t *perm, double *rec, double ro,
                     double co, double gamma, double *p, double *cs);
int sph_tait_records_f32(sph_ctx *ctx, int64_t n, const int32_t *perm, float *rec, double ro,
                         double co, double gamma, double *p, double *cs);

// P/rho^2 into SPH_R_PRHO2 of the first ns[s] records of every source
// (float records when f32)
int sph_prho2_sources(sph_ctx *ctx, const sph_sources *src, int f32);

// Exclusive prefix sums (device-wide), stream-ordered.  `data` has n
// elements; the scanned values are written in place and the total is
// written to data[n] (so data must hold n+1 elements).
int sph_scan_exclusive_i32(sph_ctx *ctx, int32_t *data, int64_t n);
int sph_scan_exclusive_i64(sph_ctx *ctx, int64_t *data, int64_t n);
// the same with the scan's partial sums placed in the context scratch from
// byte offset scratch_off on (data itself may live below that offset)
int sph_scan_exclusive_i32_at(sph_ctx *ctx, int32_t *data, int64_t n, size_t scratch_off);
int sph_scan_exclusive_i64_at(sph_ctx *ctx, int64_t *data, int64_t n, size_t scratch_off);
size_t sph_scan_scratch_bytes(int64_t n, size_t elem);

// ------------------------------------------------------------------------
// Cubic spline kernel in the reference's expression order.
// base/kernels.pyx:51-57 (h_dim), :355-378 (function), :380-417 (gradient),
// :441-447 (_fac, taken at h = 1.0 in KernelBase.__init__, :75).
// Compiled with -ffp-contract=off: no FMA contraction, IEEE sqrt and
// division, so each value is the same double the reference computes.
// ------------------------------------------------------------------------
#define SPH_PI 3.141592653589793

// ------------------------------------------------------------------------
// pow(x, e) of the equation of state and the artificial pressure
// (sph/funcs/eos_funcs.pyx:118-129, pressure_funcs.pyx:121-136).  For an
// integral exponent 1 <= e <= 32 and 2^-20 <= |x| <= 2^20 the value is x^e
// correctly rounded: the power is taken in double-double arithmetic (each
// product split exactly with fma, Dekker/Knuth two-product and fast
// two-sum), relative error below 2^-98 before the one final rounding, so the
// result is the nearest double to the exact power unless that power lies
// within 2^-98 of a rounding midpoint.  A correctly rounded pow is what the
// C library of the reference's time computed (glibc's pow before 2.28 was
// correctly rounded); device pow() and current glibc differ from it by an
// ulp now and then.  Other exponents and ranges: pow().
// ------------------------------------------------------------------------
struct sph_dd {
    double hi, lo;
};

__device__ __forceinline__ sph_dd sph_dd_mul(sph_dd a, sph_dd b)
{
    const double p = a.hi * b.hi;
    const double e = fma(a.hi, b.hi, -p);  // a.hi * b.hi == p + e exactly
    const double t = a.hi * b.lo + a.lo * b.hi;
    const double s = e + t;
    const double hi = p + s;
    return {hi, s - (hi - p)};
}

__device__ __forceinline__ double sph_pow(double x, double e)
{
    const double ax = fabs(x);
    if (!(e >= 1.0 && e <= 32.0 && e == floor(e) && ax >= 0x1p-20 && ax <= 0x1p20))
        return pow(x, e);
    const int n = (int)e;
    const sph_dd xd = {x, 0.0};
    sph_dd r = xd;
    for (int b = 30 - __builtin_clz((unsigned)n); b >= 0; --b) {  // bits below the top one
        r = sph_dd_mul(r, r);
        if ((n >> b) & 1) r = sph_dd_mul(r, xd);
    }
    return r.hi;
}

// float records: the single-precision pow
__device__ __forceinline__ float sph_pow(float x, float e) { return pow(x, e); }

// a / b correctly rounded, from inv_b = RN(1/b), the reciprocal one IEEE
// division gave (0 when b lies outside [2^-400, 2^400]): q0 = RN(a * inv_b) is
// within an ulp of a/b, the residual a - b*q0 is exact by fma, and one fma
// correction rounds it correctly (Markstein's theorem), so the result is the
// IEEE quotient bit for bit -- three operations instead of a division.  Used
// where one destination divides every pair by its own h.  |a| outside
// [2^-500, 2^500] (zero excepted) takes the IEEE division.
__device__ __forceinline__ double sph_div_inv(double a, double b, double inv_b)
{
    const double aa = fabs(a);
    if (!(inv_b != 0.0 && ((aa >= 0x1p-500 && aa <= 0x1p500) || a == 0.0))) return a / b;
    const double q0 = a * inv_b;
    return fma(fma(-b, q0, a), inv_b, q0);
}
__device__ __forceinline__ float sph_div_inv(float a, float b, float) { return a / b; }

// the reciprocal sph_div_inv takes: 1/b by IEEE division, 0 when b is out of range
template <typename T> __device__ __forceinline__ T sph_inv_for_div(T b)
{
    return (b >= T(0x1p-400) && b <= T(0x1p400)) ? T(1.0) / b : T(0.0);
}

// Every helper is a template over the arithmetic type T (double: the
// reference's FP64 path; float: the single-precision variant); the double
// instantiation evaluates exactly the expressions above.
template <typename T> __device__ __forceinline__ T sph_h_dim(T h, int dim)
{
    if (dim == 1) return T(1) / h;
    else if (dim == 2) return T(1) / (h * h);
    else return T(1) / (h * h * h);
}

template <typename T = double> __host__ __device__ __forceinline__ T sph_cubic_fac1(int dim)
{
    if (dim == 1) return T((2. / 3.) / 1.0);
    else if (dim == 2) return T(10 / (7 * SPH_PI) / (1.0 * 1.0));
    else return T(1. / SPH_PI / (1.0 * 1.0 * 1.0));
}

// W(pa, pb, h) given the three coordinate differences pa - pb
__device__ __forceinline__ double sph_cubic_w(int dim, double rx, double ry, double rz, double h)
{
    double rab = sqrt(rx * rx + ry * ry + rz * rz);
    double q = rab / h;
    double val;
    double fac = sph_h_dim(h, dim) * sph_cubic_fac1(dim);
    if (q > 2.0)
        val = 0.0;
    else if (q > 1.0)
        val = 0.25 * (2 - q) * (2 - q) * (2 - q);
    else
        val = 1 - 1.5 * (q * q) * (1 - 0.5 * q);
    return val * fac;
}

// W with the factor h_dim(h)*sigma supplied by the caller (the same value
// sph_cubic_w computes)
template <typename T>
__device__ __forceinline__ T sph_cubic_w_fac(T rx, T ry, T rz, T h, T fac, T inv_h = T(0.0))
{
    T rab = sqrt(rx * rx + ry * ry + rz * rz);
    T q = sph_div_inv(rab, h, inv_h);
    T val;
    if (q > T(2.0))
        val = T(0.0);
    else if (q > T(1.0))
        val = T(0.25) * (T(2) - q) * (T(2) - q) * (T(2) - q);
    else
        val = T(1) - T(1.5) * (q * q) * (T(1) - T(0.5) * q);
    return val * fac;
}

// gradient factor with h_dim(h)*sigma supplied by the caller.  The two
// branches of kernels.pyx:408-411 divide different numerators by different
// denominators; selecting both first and dividing once gives the same
// quotient with one division per pair instead of two divergent ones.
template <typename T>
__device__ __forceinline__ T sph_cubic_gfac_fac(T rx, T ry, T rz, T h, T fac, T inv_h = T(0.0))
{
    const T rab = sqrt(rx * rx + ry * ry + rz * rz);
    const T q = sph_div_inv(rab, h, inv_h);
    const bool outer = q >= T(1.0);
    const T num = outer ? T(-0.75) * (T(2) - q) * (T(2) - q) : T(3.0) * (T(0.75) * q - T(1));
    const T den = outer ? h * rab : h * h;
    const T v = num / den;
    const T val = (q > T(2.0) || !(outer || q > T(1e-14))) ? T(0.0) : v;
    return val * fac;
}

// gradient factor val*fac so that gradW = r * (val*fac)
__device__ __forceinline__ double sph_cubic_gfac(int dim, double rx, double ry, double rz,
                                                 double h)
{
    double rab = sqrt(rx * rx + ry * ry + rz * rz);
    double q = rab / h;
    double val = 0.0;
    double fac = sph_h_dim(h, dim) * sph_cubic_fac1(dim);
    if (q > 2.0) {
    } else if (q >= 1.0) {
        val = -0.75 * (2 - q) * (2 - q) / (h * rab);
    } else if (q > 1e-14) {
        val = 3.0 * (0.75 * q - 1) / (h * h);
    }
    return val * fac;
}

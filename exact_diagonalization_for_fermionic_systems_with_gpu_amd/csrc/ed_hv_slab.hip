// ed_hv_slab.hip -- H.v on the slab layout, the vector layout the Lanczos solver keeps
// internally for real sectors of a million amplitudes and more.
//
// Same operator and per-element summation order as Hubbard_mv_ (MBL_test/include/
// HubMatrix_GPU.cuh:114-161) and as the on-the-fly kernels of ed_hv_otf.hip: UP entries in
// H_TBq call order, then DN entries in call order, then the interaction -- the result is
// bit-identical to ed_hubbard_apply on the same amplitudes.  What changes is where the
// amplitudes live:
//
//   layout L (the solver's vectors): both species' basis states are renumbered in reverse
//     Cuthill-McKee order of their hopping graphs (positions pu, pd), and the state is cut
//     into slabs of 16 dn positions: amplitude (u, d) sits at
//         ((pd >> 4) * dim_up + pu) * 16 + (pd & 15).
//     One slab is dim_up lines of 128 B, contiguous (6.2 MB at 18 sites), so every UP gather of
//     a pass over one slab lands in a few 2 MB pages, and hops in RCM numbering stay within a
//     band of rows (+-11.4 k at 18 sites), which keeps one slab's gathers inside one XCD's L2.
//   layout T (scratch): the same with the roles exchanged, slabs of 16 up positions:
//         ((pu >> 4) * dim_dn + pd) * 16 + (pu & 15).
//
// Pass A (UP): tiles of TR (32 or 64, the autotuner's choice) up positions x one slab, slab-major,
//   slabs split over the 8 XCDs.
//   A lane owns 2 adjacent columns (16 B), 8 lanes a row, so one wavefront load fetches 8 full
//   128 B lines (the j-th neighbour row of 8 rows).  It writes the UP partial in layout L and
//   the tile's own (scaled) amplitudes, transposed through LDS, in layout T.
// Pass B (DN + diagonal + Lanczos epilogue): the DN term is an UP-type pass over layout T: tiles
//   of TR dn positions x one slab of 16 up positions.  It reads the UP partial (and the Lanczos
//   w) tile in layout L transposed through LDS, adds the DN gathers and the interaction, and
//   writes y (and v) back in layout L.
// Gathers are 16 B per lane and hit L2 instead of MALL/HBM with TLB misses, at the price of one
// extra write and read of a vector per H.v (the transposed copy).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ed_common.hpp"
#include "ed_hubbard.hpp"

namespace ed {

constexpr int kSW = 16;     // positions per slab: one 128 B line of a row
// tile rows TR (32 or 64; 8 lanes x 2 columns per row, so 8 TR threads) and gathers in flight
// per lane NB (8 or 16) are template parameters chosen by the slab autotuner
constexpr int kMaxOps = 64; // calls per species (the on-the-fly limit)

// Width of the last slab of a species with dim positions: the remainder of dim / 16 rounded up to
// even (16 B loads stay aligned), 16 when dim is a multiple of 16.  Every other slab is 16 wide,
// so a layout vector has slab_len_of(rows, dim) = rows * (dim rounded up to that width) entries:
// exactly rows * dim when dim is even.
static int last_slab_width(int64_t dim)
{
    const int r = (int)(dim - ((dim + kSW - 1) / kSW - 1) * kSW);
    return r + (r & 1);
}
static int64_t slab_len_of(int64_t rows, int64_t dim)
{
    return ((dim + kSW - 1) / kSW - 1) * rows * kSW + rows * last_slab_width(dim);
}

struct SlabLayout {
    int64_t du = 0, dd = 0, nsd = 0, nsu = 0, len = 0, xt_len = 0;
    int nops_u = 0, nops_d = 0;
    int32_t *pos_u = nullptr, *pos_d = nullptr; // rank -> position
    int32_t *u_at = nullptr, *d_at = nullptr;   // position -> rank
    int32_t *st_u = nullptr, *st_d = nullptr;   // basis state by position
    int32_t *ptr_u = nullptr, *ptr_d = nullptr; // CSR over positions, entries in call order
    uint32_t *ent_u = nullptr, *ent_d = nullptr; // (neighbour position << 8) | value code
    double *val_u = nullptr, *val_d = nullptr;  // value code -> matrix element (4 per call)
    double *xt = nullptr;                       // layout-T scratch
    double *part = nullptr;                     // pass-B block partials of <v, y>
    int64_t npart = 0;                          // pass-B tiles at the smallest tile (32 rows)
    int tr = 32, nb = 8, swz = 8;               // launch shape: tile rows, gathers in flight, tile order
    int wl = kSW, wt = kSW;                     // widths of the last slab of layout L and of layout T
};

void slab_free(SlabLayout *s)
{
    if (!s) return;
    for (void *p : {(void *)s->pos_u, (void *)s->pos_d, (void *)s->u_at, (void *)s->d_at, (void *)s->st_u,
                    (void *)s->st_d, (void *)s->ptr_u, (void *)s->ptr_d, (void *)s->ent_u, (void *)s->ent_d,
                    (void *)s->val_u, (void *)s->val_d, (void *)s->xt, (void *)s->part})
        dfree(p);
    delete s;
}

struct SlabArgs {
    int64_t du, dd, nsd, nsu;
    const int32_t *ptr;   // CSR of the species the pass applies
    const uint32_t *ent;
    const double *val;
    int nops;
    const int32_t *st_u, *st_d;
    double U;
    int64_t mask;
    const double *Uvec;
    int has_diag;
    int swz;    // tiles per CU slot in the tile order (slab_tile); 0 or 1: plain order
    int wl, wt; // line widths of the last slab of layout L (dn) and of layout T (up); others 16
};

// Tile of workgroup b.  The 8 XCDs take contiguous eighths of the tiles (workgroups are dealt to
// the XCDs round-robin).  With swz > 1, each run of 32 * swz consecutive workgroups of one XCD
// takes its tiles transposed: workgroups i and i + 32 (the next one dealt to the same CU when
// dispatch goes round-robin over the XCD's 32 CUs) get adjacent tiles, so that the workgroups
// resident on one CU gather from overlapping rows.  A bijection on every full run; the last
// partial run keeps the plain order.
__device__ __forceinline__ int64_t slab_tile(int64_t b, int64_t per_xcd, int swz)
{
    int64_t i = b >> 3;
    if (swz > 1) {
        const int64_t G = 32 * (int64_t)swz, g = i / G, r = i - g * G;
        if ((g + 1) * G <= per_xcd) i = g * G + (r & 31) * swz + (r >> 5);
    }
    return (b & 7) * per_xcd + i;
}

__device__ __forceinline__ double slab_diag(int32_t su, int32_t sd, const SlabArgs &a)
{
    const uint32_t dbl = (uint32_t)(su & sd) & (uint32_t)a.mask;
    if (!a.Uvec) return a.U * (double)__popc(dbl);
    double s = 0.0;
    uint32_t m = dbl;
    while (m) {
        const int i = __ffs(m) - 1;
        s += a.Uvec[i];
        m &= m - 1;
    }
    return s;
}

// Streaming stores and loads of the layout-L / layout-T vectors are nontemporal, so they do not
// evict the gather window (the +-band of rows around the tiles in flight) from L2.
typedef double d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void st2(double *p, double a, double b)
{
    d2v v = {a, b};
    __builtin_nontemporal_store(v, reinterpret_cast<d2v *>(p));
}
__device__ __forceinline__ double2 ld2(const double *p)
{
    const d2v v = __builtin_nontemporal_load(reinterpret_cast<const d2v *>(p));
    return make_double2(v.x, v.y);
}

// LDS staging shared by both passes: value table and the CSR entries of rows [p0, pend)
template <int TR>
__device__ __forceinline__ int slab_stage(const SlabArgs &a, int64_t p0, int64_t pend, double *s_val,
                                          uint32_t *s_ent, int *s_off)
{
    constexpr int NT = 8 * TR;
    const int tid = threadIdx.x;
    for (int q = tid; q < 4 * a.nops; q += NT) s_val[q] = a.val[q];
    const int nrows = (int)(pend - p0);
    const int64_t e0 = a.ptr[p0];
    if (tid <= TR) s_off[tid] = (int)(a.ptr[tid <= nrows ? p0 + tid : pend] - e0);
    const int ne = (int)(a.ptr[pend] - e0);
    for (int q = tid; q < ne; q += NT) s_ent[q] = a.ent[e0 + q];
    return nrows;
}

// Sum over the row's entries (call order) of value * row[neighbour][c..c+1]; x scaled by sc.
// The gathers of NB entries are issued before their FMAs (NB loads of 16 B in flight per lane).
template <bool SCALE, int NB>
__device__ __forceinline__ void slab_gather(const double *__restrict__ base, int W, const uint32_t *s_ent,
                                            const double *s_val, int o0, int o1, int c, double sc, double &acc0,
                                            double &acc1)
{
    for (int j0 = o0; j0 < o1; j0 += NB) {
        double2 xv[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k)
            if (j0 + k < o1) xv[k] = *reinterpret_cast<const double2 *>(base + (int64_t)(s_ent[j0 + k] >> 8) * W + c);
#pragma unroll
        for (int k = 0; k < NB; ++k)
            if (j0 + k < o1) {
                const double av = s_val[s_ent[j0 + k] & 0xffu];
                double x0 = xv[k].x, x1 = xv[k].y;
                if (SCALE) {
                    x0 = sc * x0;
                    x1 = sc * x1;
                }
                acc0 = fma(av, x0, acc0);
                acc1 = fma(av, x1, acc1);
            }
    }
}

// Pass A: z = UP term of x (layout L); xt = sc*x (layout T)
template <bool SCALE, int NB, int TR>
__global__ __launch_bounds__(8 * TR) void k_slab_up(SlabArgs a, int64_t ntiles, int64_t per_xcd, int64_t ntr,
                                                 const double *__restrict__ x, double *__restrict__ z,
                                                 double *__restrict__ xt, const double *__restrict__ bdev, int step)
{
    __shared__ double s_val[4 * kMaxOps];
    __shared__ uint32_t s_ent[TR * kMaxOps];
    __shared__ int s_off[TR + 1];
    __shared__ double s_x[TR][kSW + 1];
    const int64_t b = blockIdx.x;
    const int64_t t = slab_tile(b, per_xcd, a.swz);
    if (t >= ntiles) return; // uniform over the workgroup
    const int64_t slab = t / ntr, p0 = (t - slab * ntr) * TR;
    const int64_t pend = (p0 + TR < a.du) ? p0 + TR : a.du;
    const int nrows = slab_stage<TR>(a, p0, pend, s_val, s_ent, s_off);
    __syncthreads();
    const double sc = SCALE ? 1.0 / bdev[step] : 1.0;
    const int tid = threadIdx.x, r = tid >> 3, c = (tid & 7) * 2;
    const double *xs = x + slab * a.du * kSW;
    const int W = slab == a.nsd - 1 ? a.wl : kSW; // line width of this slab (lanes c >= W idle)
    if (r < nrows && c < W) {
        const int64_t p = p0 + r;
        double acc0 = 0.0, acc1 = 0.0;
        slab_gather<SCALE, NB>(xs, W, s_ent, s_val, s_off[r], s_off[r + 1], c, sc, acc0, acc1);
        st2(z + slab * a.du * kSW + p * W + c, acc0, acc1);
        double2 xo = *reinterpret_cast<const double2 *>(xs + p * W + c);
        if (SCALE) {
            xo.x = sc * xo.x;
            xo.y = sc * xo.y;
        }
        s_x[r][c] = xo.x;
        s_x[r][c + 1] = xo.y;
    } else {
        s_x[r][c] = 0.0;
        s_x[r][c + 1] = 0.0;
    }
    __syncthreads();
    // transposed copy: (pu = p0 + 16 us + k, pd = 16 slab + cc) -> layout T,
    // (pu >> 4) dd 16 + pd Wt + (pu & 15) with Wt the width of up slab pu >> 4
    const int us = tid >> 7, rem = tid & 127, cc = rem >> 3, k = (rem & 7) * 2;
    const int64_t pd = slab * kSW + cc, pu0 = p0 + us * kSW;
    const int Wt = (pu0 >> 4) == a.nsu - 1 ? a.wt : kSW;
    if (pd < a.dd && pu0 < a.du && k < Wt)
        st2(xt + (pu0 >> 4) * a.dd * kSW + pd * Wt + k, s_x[us * kSW + k][cc], s_x[us * kSW + k + 1][cc]);
}

// Pass B option of the Ritz-vector rebuild, where a[step] is already known: y is written as
// f = y - a[step] v and evec += zr v, the arithmetic of ritz_update (ed_blas.hip) element for
// element, so the rebuilt Lanczos vectors are the forward recurrence's bit for bit.
enum { SLAB_RITZ = 16 };
// Pass B option of the recurrence that keeps no v: w is read as f_{N-1} and scaled by
// 1 / b[step - 1] on the fly -- the product pass A formed when it scaled x, so w is the v the
// previous step would have written, bit for bit.
enum { SLAB_WSCALE = 32 };

// Pass B: y = zin + DN term of xt + diag * xt (+ c w); v = xt (layout L); block partial <v, y>
template <int FLAGS, int NB, int TR>
__global__ __launch_bounds__(8 * TR) void k_slab_dn(SlabArgs a, int64_t ntiles, int64_t per_xcd, int64_t ntr,
                                                 const double *__restrict__ xt, const double *zin, const double *win,
                                                 double *yout, double *vout, const double *__restrict__ bdev, int step,
                                                 double *__restrict__ partial, double *__restrict__ evec,
                                                 const double *__restrict__ adev, double zr)
{
    __shared__ double s_val[4 * kMaxOps];
    __shared__ uint32_t s_ent[TR * kMaxOps];
    __shared__ int s_off[TR + 1];
    __shared__ double s_a[TR][kSW + 1];
    __shared__ double s_b[TR][kSW + 1];
    __shared__ double s_red[TR / 8];
    const int64_t b = blockIdx.x;
    const int64_t t = slab_tile(b, per_xcd, a.swz);
    if (t >= ntiles) return; // uniform over the workgroup
    const int64_t su = t / ntr, p0 = (t - su * ntr) * TR;
    const int64_t pend = (p0 + TR < a.dd) ? p0 + TR : a.dd;
    const int nrows = slab_stage<TR>(a, p0, pend, s_val, s_ent, s_off);
    const int tid = threadIdx.x;
    // the tile's UP partial (and w) in layout L: TR/16 dn slabs x 16 up positions; a line of dn
    // slab dsl has Wl entries (the last slab may be narrower)
    const int ds = tid >> 7, rem = tid & 127, pr = rem >> 3, k = (rem & 7) * 2;
    const int64_t pul = su * kSW + pr, dsl = (p0 >> 4) + ds;
    const int Wl = dsl == a.nsd - 1 ? a.wl : kSW;
    const bool own = pul < a.du && dsl < a.nsd && k < Wl;
    const int64_t li = dsl * a.du * kSW + pul * Wl + k;
    {
        double2 zz = make_double2(0.0, 0.0), ww = make_double2(0.0, 0.0);
        if (own) {
            zz = ld2(zin + li);
            if (FLAGS & HV_AXPY) ww = ld2(win + li);
            if (FLAGS & SLAB_WSCALE) {
                const double sw = 1.0 / bdev[step - 1];
                ww.x = sw * ww.x;
                ww.y = sw * ww.y;
            }
        }
        s_a[ds * kSW + k][pr] = zz.x;
        s_a[ds * kSW + k + 1][pr] = zz.y;
        s_b[ds * kSW + k][pr] = ww.x;
        s_b[ds * kSW + k + 1][pr] = ww.y;
    }
    __syncthreads();
    const int r = tid >> 3, c = (tid & 7) * 2;
    const int64_t pd = p0 + r, pu = su * kSW + c;
    const int Wu = su == a.nsu - 1 ? a.wt : kSW; // line width of this up slab in layout T
    double o0 = 0.0, o1 = 0.0, xd0 = 0.0, xd1 = 0.0, dot = 0.0;
    if (r < nrows && c < Wu) {
        double acc0 = s_a[r][c], acc1 = s_a[r][c + 1];
        const double *xs = xt + su * a.dd * kSW;
        slab_gather<false, NB>(xs, Wu, s_ent, s_val, s_off[r], s_off[r + 1], c, 1.0, acc0, acc1);
        const double2 xo = *reinterpret_cast<const double2 *>(xs + pd * Wu + c);
        xd0 = xo.x;
        xd1 = xo.y;
        if (a.has_diag) {
            const int32_t sdn = a.st_d[pd];
            if (pu < a.du) acc0 = fma(slab_diag(a.st_u[pu], sdn, a), xd0, acc0);
            if (pu + 1 < a.du) acc1 = fma(slab_diag(a.st_u[pu + 1], sdn, a), xd1, acc1);
        }
        o0 = acc0;
        o1 = acc1;
        if (FLAGS & HV_AXPY) {
            const double cw = -bdev[step];
            o0 = fma(cw, s_b[r][c], o0);
            o1 = fma(cw, s_b[r][c + 1], o1);
        }
        if (pu >= a.du) o0 = xd0 = 0.0; // padding columns of the last up slab
        if (pu + 1 >= a.du) o1 = xd1 = 0.0;
        if (FLAGS & HV_DOT) {
            dot = xd0 * o0;
            dot = fma(xd1, o1, dot);
        }
    }
    __syncthreads(); // s_a / s_b read by every lane
    s_a[r][c] = o0;
    s_a[r][c + 1] = o1;
    s_b[r][c] = xd0;
    s_b[r][c + 1] = xd1;
    __syncthreads();
    if (own) {
        double y0 = s_a[ds * kSW + k][pr], y1 = s_a[ds * kSW + k + 1][pr];
        const double v0 = s_b[ds * kSW + k][pr], v1 = s_b[ds * kSW + k + 1][pr];
        if (FLAGS & SLAB_RITZ) {
            const double ca = -adev[step];
            y0 = fma(ca, v0, y0);
            y1 = fma(ca, v1, y1);
            const double2 e = ld2(evec + li);
            st2(evec + li, fma(zr, v0, e.x), fma(zr, v1, e.y));
        }
        st2(yout + li, y0, y1);
        if (FLAGS & HV_WRITEV) st2(vout + li, v0, v1);
    }
    if (FLAGS & HV_DOT) {
        // one partial per 32-row unit (its four waves summed in order), at slot su ntr32 + unit:
        // the same partials, in the same slots, for every tile height
        const double ws = wave_sum(dot);
        if ((tid & (kWave - 1)) == 0) s_red[tid >> 6] = ws;
        __syncthreads();
        const int64_t ntr32 = (a.dd + 31) / 32, unit = (p0 >> 5) + tid;
        if (tid < TR / 32 && unit < ntr32) {
            const double *w4 = s_red + 4 * tid;
            partial[su * ntr32 + unit] = ((w4[0] + w4[1]) + w4[2]) + w4[3];
        }
    }
}

// Gather ceiling: the line gathers of one pass alone.  Same grid, tile order, LDS footprint,
// entry staging and per-lane gather batches as k_slab_up (UP entries over layout L) or k_slab_dn
// (DN entries over layout T), but no value table, no FMAs and no stream loads or stores: the
// loaded amplitudes are only summed, and the sum is stored under a condition that real data
// never meets, so that the loads stay live.  Its time is what the gathers of that pass cost.
template <int NB, int TR>
__global__ __launch_bounds__(8 * TR) void k_slab_ceiling(SlabArgs a, int64_t ntiles, int64_t per_xcd, int64_t ntr,
                                                         int64_t drows, int64_t nslabs, int wlast,
                                                         const double *__restrict__ x, double *__restrict__ sink,
                                                         int64_t nsink)
{
    __shared__ double s_val[4 * kMaxOps];
    __shared__ uint32_t s_ent[TR * kMaxOps];
    __shared__ int s_off[TR + 1];
    __shared__ double s_a[TR][kSW + 1];
    __shared__ double s_b[TR][kSW + 1];
    const int64_t b = blockIdx.x;
    const int64_t t = slab_tile(b, per_xcd, a.swz);
    if (t >= ntiles) return; // uniform over the workgroup
    const int64_t slab = t / ntr, p0 = (t - slab * ntr) * TR;
    const int64_t pend = (p0 + TR < drows) ? p0 + TR : drows;
    const int nrows = slab_stage<TR>(a, p0, pend, s_val, s_ent, s_off);
    __syncthreads();
    const int tid = threadIdx.x, r = tid >> 3, c = (tid & 7) * 2;
    const double *xs = x + slab * drows * kSW;
    const int W = slab == nslabs - 1 ? wlast : kSW;
    double acc = 0.0;
    if (r < nrows && c < W) {
        const int o0 = s_off[r], o1 = s_off[r + 1];
        for (int j0 = o0; j0 < o1; j0 += NB) {
            double2 xv[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k)
                if (j0 + k < o1) xv[k] = *reinterpret_cast<const double2 *>(xs + (int64_t)(s_ent[j0 + k] >> 8) * W + c);
#pragma unroll
            for (int k = 0; k < NB; ++k)
                if (j0 + k < o1) acc += xv[k].x + xv[k].y;
        }
    }
    s_a[r][c] = acc;      // the LDS arrays of the real passes are declared (same occupancy) and
    s_b[r][c] = s_val[0]; // touched so that the compiler keeps them
    if (acc == -1.0e300 && t < nsink) sink[t] = s_a[r][c] + s_b[r][c];
}

// fixed-order sums of 2048 consecutive partials (first level of the <v, y> reduction)
__global__ __launch_bounds__(256) void k_slab_chunks(int64_t n, const double *__restrict__ in, double *__restrict__ out)
{
    __shared__ double red[4];
    const int64_t base = (int64_t)blockIdx.x * 2048;
    double s = 0.0;
    for (int q = 0; q < 8; ++q) {
        const int64_t i = base + q * 256 + threadIdx.x;
        if (i < n) s += in[i];
    }
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// reference layout -> layout L (the narrow last slab's padding column, if any, written zero)
__global__ __launch_bounds__(256) void k_slab_from_rm(SlabArgs a, const int32_t *__restrict__ u_at,
                                                      const int32_t *__restrict__ d_at, const double *__restrict__ x,
                                                      double *__restrict__ xl, int64_t len)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= len) return;
    int64_t slab = i / (a.du * kSW);
    if (slab > a.nsd - 1) slab = a.nsd - 1;
    const int W = slab == a.nsd - 1 ? a.wl : kSW;
    const int64_t rem = i - slab * a.du * kSW, p = rem / W, pd = slab * kSW + (rem - p * W);
    xl[i] = pd < a.dd ? x[(int64_t)u_at[p] * a.dd + d_at[pd]] : 0.0;
}

__global__ __launch_bounds__(256) void k_slab_to_rm(SlabArgs a, const int32_t *__restrict__ pos_u,
                                                    const int32_t *__restrict__ pos_d, const double *__restrict__ xl,
                                                    double *__restrict__ x)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= a.du * a.dd) return;
    const int64_t u = i / a.dd, d = i - u * a.dd;
    const int64_t pd = pos_d[d], slab = pd >> 4;
    const int W = slab == a.nsd - 1 ? a.wl : kSW;
    x[i] = xl[slab * a.du * kSW + (int64_t)pos_u[u] * W + (pd & 15)];
}

static SlabArgs slab_args(const Hubbard &h, bool up)
{
    const SlabLayout &L = *h.slab;
    SlabArgs a;
    a.du = L.du;
    a.dd = L.dd;
    a.nsd = L.nsd;
    a.nsu = L.nsu;
    a.ptr = up ? L.ptr_u : L.ptr_d;
    a.ent = up ? L.ent_u : L.ent_d;
    a.val = up ? L.val_u : L.val_d;
    a.nops = up ? L.nops_u : L.nops_d;
    a.st_u = L.st_u;
    a.st_d = L.st_d;
    a.U = h.otf.U;
    a.mask = h.otf.mask;
    a.Uvec = h.otf.Uvec;
    a.has_diag = h.otf.has_diag;
    a.swz = L.swz;
    a.wl = L.wl;
    a.wt = L.wt;
    return a;
}

int64_t slab_len(const Hubbard &h) { return h.slab ? h.slab->len : 0; }
int64_t slab_len_for(const Hubbard &h) { return slab_len_of(h.otf.dim_up, h.otf.dim_dn); }
int64_t slab_num_partials(const Hubbard &h) { return h.slab ? (h.slab->npart + 2047) / 2048 : 0; }

template <int NB, int TR>
static int slab_apply_t(const Hubbard &h, const double *x, double *z, const double *w, double *y, double *v,
                        const double *bdev, int step, int flags, double *partial, hipStream_t st, int64_t *nblocks,
                        double *evec, const double *adev, double zr)
{
    const SlabLayout &L = *h.slab;
    SlabArgs au = slab_args(h, true), ad = slab_args(h, false);
    const int64_t ntrA = (L.du + TR - 1) / TR, tilesA = L.nsd * ntrA, perA = (tilesA + 7) / 8;
    const int64_t ntrB = (L.dd + TR - 1) / TR, tilesB = L.nsu * ntrB, perB = (tilesB + 7) / 8;
    if (perA * 8 > 0x7fffffffLL || perB * 8 > 0x7fffffffLL) return fail(ED_EUNSUPPORTED, "slab H.v: grid too large");
    const dim3 gA((unsigned)(perA * 8)), gB((unsigned)(perB * 8)), blk(8 * TR);
    if (flags & HV_SCALE)
        hipLaunchKernelGGL((k_slab_up<true, NB, TR>), gA, blk, 0, st, au, tilesA, perA, ntrA, x, z, L.xt, bdev, step);
    else
        hipLaunchKernelGGL((k_slab_up<false, NB, TR>), gA, blk, 0, st, au, tilesA, perA, ntrA, x, z, L.xt, bdev, step);
    ED_CHECK_LAUNCH();
    const int fb = flags & ~HV_SCALE;
#define ED_SLAB_B(F)                                                                                               \
    hipLaunchKernelGGL((k_slab_dn<F, NB, TR>), gB, blk, 0, st, ad, tilesB, perB, ntrB, L.xt, z, w, y, v, bdev, step, \
                       L.part, evec, adev, zr)
    switch (fb) {
    case 0: ED_SLAB_B(0); break;
    case HV_WRITEV: ED_SLAB_B(HV_WRITEV); break;
    case HV_WRITEV | HV_AXPY: ED_SLAB_B(HV_WRITEV | HV_AXPY); break;
    case HV_WRITEV | HV_DOT: ED_SLAB_B(HV_WRITEV | HV_DOT); break;
    case HV_WRITEV | HV_AXPY | HV_DOT: ED_SLAB_B(HV_WRITEV | HV_AXPY | HV_DOT); break;
    case HV_WRITEV | SLAB_RITZ: ED_SLAB_B(HV_WRITEV | SLAB_RITZ); break;
    case HV_WRITEV | HV_AXPY | SLAB_RITZ: ED_SLAB_B(HV_WRITEV | HV_AXPY | SLAB_RITZ); break;
    case HV_DOT: ED_SLAB_B(HV_DOT); break;
    case HV_AXPY | HV_DOT | SLAB_WSCALE: ED_SLAB_B(HV_AXPY | HV_DOT | SLAB_WSCALE); break;
    case SLAB_RITZ: ED_SLAB_B(SLAB_RITZ); break;
    case HV_AXPY | SLAB_RITZ | SLAB_WSCALE: ED_SLAB_B(HV_AXPY | SLAB_RITZ | SLAB_WSCALE); break;
    default: return fail(ED_EINVAL, "slab H.v: unsupported flags %d", flags);
    }
#undef ED_SLAB_B
    ED_CHECK_LAUNCH();
    if (fb & HV_DOT) {
        const int64_t np32 = L.nsu * ((L.dd + 31) / 32); // pass-B partials, one per 32-row unit
        const int64_t nc = (np32 + 2047) / 2048;
        hipLaunchKernelGGL(k_slab_chunks, dim3((unsigned)nc), dim3(256), 0, st, np32, L.part, partial);
        ED_CHECK_LAUNCH();
        if (nblocks) *nblocks = nc;
    }
    return ED_OK;
}

// one H.v on layout L with the layout's launch shape; flags as hv_lanczos (HV_SCALE applies to
// pass A, which pre-scales xt)
static int slab_apply(const Hubbard &h, const double *x, double *z, const double *w, double *y, double *v,
                      const double *bdev, int step, int flags, double *partial, hipStream_t st, int64_t *nblocks,
                      double *evec = nullptr, const double *adev = nullptr, double zr = 0.0)
{
    const SlabLayout &L = *h.slab;
#define ED_SLAB_T(NB, TR) slab_apply_t<NB, TR>(h, x, z, w, y, v, bdev, step, flags, partial, st, nblocks, evec, adev, zr)
    if (L.tr == 64) return L.nb == 16 ? ED_SLAB_T(16, 64) : ED_SLAB_T(8, 64);
    return L.nb == 16 ? ED_SLAB_T(16, 32) : ED_SLAB_T(8, 32);
#undef ED_SLAB_T
}

int slab_lanczos(const Hubbard &h, const double *f, double *y, double *v, const double *w, const double *bdev,
                 int step, bool dot, double *partial, hipStream_t st, int64_t *nblocks)
{
    int flags = HV_SCALE | HV_WRITEV | (step > 0 ? HV_AXPY : 0) | (dot ? HV_DOT : 0);
    // the UP partial goes into v (overwritten by pass B with v itself), w and y share a buffer
    return slab_apply(h, f, v, w, y, v, bdev, step, flags, partial, st, nblocks);
}

// The recurrence without v (f_{N-1} stands for v_{N-1}): x = f / b[step] (layout L) -> y = H v -
// b[step] f_prev / b[step - 1] in y, which also holds the UP partial; f_prev is not written.
int slab_lanczos_vfree(const Hubbard &h, const double *f, const double *f_prev, double *y, const double *bdev,
                       int step, double *partial, hipStream_t st, int64_t *nblocks)
{
    const int flags = HV_SCALE | HV_DOT | (step > 0 ? HV_AXPY | SLAB_WSCALE : 0);
    return slab_apply(h, f, y, f_prev, y, nullptr, bdev, step, flags, partial, st, nblocks);
}

// one step of the Ritz-vector rebuild without v: f_{i+1} = y - a[i] v_i lands in y, evec += zr v_i
int slab_lanczos_ritz_vfree(const Hubbard &h, const double *f, const double *f_prev, double *y, const double *bdev,
                            const double *adev, double zr, double *evec, int step, hipStream_t st)
{
    if (!evec || !adev) return fail(ED_EINVAL, "slab Ritz step: null vector");
    const int flags = HV_SCALE | SLAB_RITZ | (step > 0 ? HV_AXPY | SLAB_WSCALE : 0);
    return slab_apply(h, f, y, f_prev, y, nullptr, bdev, step, flags, nullptr, st, nullptr, evec, adev, zr);
}

int slab_lanczos_ritz(const Hubbard &h, const double *f, double *y, double *v, const double *bdev,
                      const double *adev, double zr, double *evec, int step, hipStream_t st)
{
    if (!evec || !adev) return fail(ED_EINVAL, "slab Ritz step: null vector");
    const int flags = HV_SCALE | HV_WRITEV | (step > 0 ? HV_AXPY : 0) | SLAB_RITZ;
    return slab_apply(h, f, v, y, y, v, bdev, step, flags, nullptr, st, nullptr, evec, adev, zr);
}

int slab_from_rowmajor(const Hubbard &h, const double *x, double *xl, hipStream_t st)
{
    const SlabLayout &L = *h.slab;
    hipLaunchKernelGGL(k_slab_from_rm, dim3((unsigned)((L.len + 255) / 256)), dim3(256), 0, st, slab_args(h, true),
                       L.u_at, L.d_at, x, xl, L.len);
    ED_CHECK_LAUNCH();
    return ED_OK;
}

int slab_to_rowmajor(const Hubbard &h, const double *xl, double *x, hipStream_t st)
{
    const SlabLayout &L = *h.slab;
    hipLaunchKernelGGL(k_slab_to_rm, dim3((unsigned)((L.du * L.dd + 255) / 256)), dim3(256), 0, st,
                       slab_args(h, true), L.pos_u, L.pos_d, xl, x);
    ED_CHECK_LAUNCH();
    return ED_OK;
}

// host: RCM positions and the per-position entry lists of one species, in call order (the order
// in which k_up_entries and the DN descriptor cache of ed_hv_otf.hip visit the calls)
static int species_tables(int N, const std::vector<OtfOp> &ops, const std::vector<int32_t> &state,
                          const std::vector<int32_t> &addr, std::vector<int32_t> &at, std::vector<int32_t> &pos,
                          std::vector<int32_t> &ptr, std::vector<uint32_t> &ent, std::vector<double> &val)
{
    const int64_t dim = (int64_t)state.size();
    const int Q = __builtin_popcount((uint32_t)state[0]);
    std::vector<ed_op> eo(ops.size());
    for (size_t o = 0; o < ops.size(); ++o) eo[o] = ed_op{ops[o].kind, ops[o].i, ops[o].j, 0, 0.0, 0.0};
    at.clear();
    if (dim > 1) ED_TRY(up_row_order(N, Q, eo.data(), (int)eo.size(), at));
    if ((int64_t)at.size() != dim) {
        at.resize((size_t)dim);
        for (int64_t r = 0; r < dim; ++r) at[(size_t)r] = (int32_t)r;
    }
    pos.assign((size_t)dim, 0);
    for (int64_t p = 0; p < dim; ++p) pos[(size_t)at[(size_t)p]] = (int32_t)p;
    ptr.assign((size_t)dim + 1, 0);
    ent.clear();
    ent.reserve((size_t)dim * 24);
    for (int64_t p = 0; p < dim; ++p) {
        const uint32_t s = (uint32_t)state[(size_t)at[(size_t)p]];
        for (size_t o = 0; o < ops.size(); ++o) {
            const OtfOp &op = ops[o];
            const uint32_t oi = (s >> op.i) & 1u, oj = (s >> op.j) & 1u;
            uint32_t e = 0xffffffffu;
            if (op.kind == ED_OP_DENSITY) {
                if (oi) e = ((uint32_t)p << 8) | (uint32_t)(4 * o);
            } else if (op.kind == ED_OP_DENSITY2) {
                if (oi & oj) e = ((uint32_t)p << 8) | (uint32_t)(4 * o);
            } else if (oi != oj) {
                const uint32_t code = (oi ? 0u : 2u) | (uint32_t)(__builtin_popcount(s & op.between) & 1);
                e = ((uint32_t)pos[(size_t)addr[s ^ op.m]] << 8) | (uint32_t)(4 * o) | code;
            }
            if (e != 0xffffffffu) ent.push_back(e);
        }
        ptr[(size_t)p + 1] = (int32_t)ent.size();
    }
    val.assign(4 * ops.size(), 0.0);
    for (size_t o = 0; o < ops.size(); ++o) {
        val[4 * o + 0] = ops[o].vf_re;
        val[4 * o + 1] = -ops[o].vf_re;
        val[4 * o + 2] = ops[o].vb_re;
        val[4 * o + 3] = -ops[o].vb_re;
    }
    return ED_OK;
}

template <typename T>
static int upload(T **dst, const std::vector<T> &src)
{
    ED_TRY(dmalloc((void **)dst, sizeof(T) * std::max<size_t>(src.size(), 1)));
    if (!src.empty() && hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice) != hipSuccess)
        return fail(ED_EHIP, "slab layout: upload failed");
    return ED_OK;
}

static int slab_build(Hubbard &h)
{
    if (h.algo != 3 || h.cplx_tables || h.has_fermi || !h.otf.up.ops_own || !h.otf.dn.ops_own)
        return fail(ED_EUNSUPPORTED, "slab layout: needs a real operator from ed_hubbard_create_ops");
    if (h.N > 26 || h.otf.up.nops > kMaxOps || h.otf.dn.nops > kMaxOps || h.otf.dim_up >= (1 << 24) ||
        h.otf.dim_dn >= (1 << 24))
        return fail(ED_EUNSUPPORTED, "slab layout: N <= 26, at most 64 calls and dimensions below 2^24");
    const int64_t du = h.otf.dim_up, dd = h.otf.dim_dn, full = (int64_t)1 << h.N;
    std::vector<int32_t> su((size_t)du), sd((size_t)dd), au((size_t)full), ad((size_t)full);
    if (hipMemcpy(su.data(), h.d_state_up, sizeof(int32_t) * du, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(sd.data(), h.d_state_dn, sizeof(int32_t) * dd, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(au.data(), h.d_addr_up, sizeof(int32_t) * full, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(ad.data(), h.d_addr_dn, sizeof(int32_t) * full, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(ED_EHIP, "slab layout: basis readback failed");
    std::vector<int32_t> atu, posu, ptru, atd, posd, ptrd;
    std::vector<uint32_t> entu, entd;
    std::vector<double> valu, vald;
    ED_TRY(species_tables(h.N, h.h_ops_up, su, au, atu, posu, ptru, entu, valu));
    ED_TRY(species_tables(h.N, h.h_ops_dn, sd, ad, atd, posd, ptrd, entd, vald));
    std::vector<int32_t> spu((size_t)du), spd((size_t)dd);
    for (int64_t p = 0; p < du; ++p) spu[(size_t)p] = su[(size_t)atu[(size_t)p]];
    for (int64_t p = 0; p < dd; ++p) spd[(size_t)p] = sd[(size_t)atd[(size_t)p]];
    SlabLayout *L = new SlabLayout;
    L->du = du;
    L->dd = dd;
    L->nsd = (dd + kSW - 1) / kSW;
    L->nsu = (du + kSW - 1) / kSW;
    L->wl = last_slab_width(dd);
    L->wt = last_slab_width(du);
    L->len = slab_len_of(du, dd);
    L->xt_len = slab_len_of(dd, du);
    L->nops_u = h.otf.up.nops;
    L->nops_d = h.otf.dn.nops;
    L->npart = L->nsu * ((dd + 31) / 32);
    int s = ED_OK;
    if ((s = upload(&L->pos_u, posu)) || (s = upload(&L->pos_d, posd)) || (s = upload(&L->u_at, atu)) ||
        (s = upload(&L->d_at, atd)) || (s = upload(&L->st_u, spu)) || (s = upload(&L->st_d, spd)) ||
        (s = upload(&L->ptr_u, ptru)) || (s = upload(&L->ptr_d, ptrd)) || (s = upload(&L->ent_u, entu)) ||
        (s = upload(&L->ent_d, entd)) || (s = upload(&L->val_u, valu)) || (s = upload(&L->val_d, vald)) ||
        (s = dmalloc((void **)&L->xt, sizeof(double) * (size_t)L->xt_len)) ||
        (s = hipMemset(L->xt, 0, sizeof(double) * (size_t)L->xt_len) == hipSuccess ? ED_OK : ED_EHIP) ||
        (s = dmalloc((void **)&L->part, sizeof(double) * (size_t)L->npart))) {
        slab_free(L);
        return s;
    }
    h.slab = L;
    return ED_OK;
}

// ed_lanczos_eigen and ed_lanczos_tridiagonal on real vectors of 2^27 amplitudes (1 GiB) and more
// build the slab layout on first use when it is supported and its scratch vector plus the nvec
// layout vectors the solver will allocate fit in free device memory (keeping 5 % of the card
// free); ED_SLAB_AUTO=0 turns this off.  The layout stays in the handle
// (ed_hubbard_slab_layout(h, 0, ...) frees it).

static bool slab_auto_fits(double free_bytes, double total_bytes, int64_t du, int64_t dd, int nvec)
{
    if ((double)du * (double)dd < (double)(1 << 27)) return false;
    return free_bytes >= (double)(nvec + 1) * 8.0 * (double)slab_len_of(du, dd) + 0.05 * total_bytes;
}

void slab_auto(Hubbard &h, int nvec)
{
    if (h.slab || h.algo != 3 || h.cplx_tables || h.has_fermi || h.N > 26) return;
    const char *e = getenv("ED_SLAB_AUTO");
    if (e && atoi(e) == 0) return;
    const int64_t du = h.otf.dim_up, dd = h.otf.dim_dn;
    if ((double)du * (double)dd < (double)(1 << 27)) return;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return;
    if (!slab_auto_fits((double)free_b, (double)total_b, du, dd, nvec)) return;
    if (slab_build(h) != ED_OK) (void)hipGetLastError(); // stays on the reference layout
}

} // namespace ed

using namespace ed;

extern "C" int ed_slab_auto_decision(double free_bytes, double total_bytes, int64_t dim_up, int64_t dim_dn, int nvec)
{
    return slab_auto_fits(free_bytes, total_bytes, dim_up, dim_dn, nvec) ? 1 : 0;
}

extern "C" int ed_hubbard_slab_layout(ed_hubbard hh, int enable, int64_t *len)
{
    Hubbard *h = reinterpret_cast<Hubbard *>(hh);
    if (!h) return fail(ED_EINVAL, "ed_hubbard_slab_layout: null operator");
    if (enable < 0) { // query
        if (len) *len = h->slab ? h->slab->len : 0;
        return ED_OK;
    }
    if (!enable) {
        ED_HIP(hipDeviceSynchronize());
        slab_free(h->slab);
        h->slab = nullptr;
        if (len) *len = 0;
        return ED_OK;
    }
    if (!h->slab) ED_TRY(slab_build(*h));
    if (len) *len = h->slab->len;
    return ED_OK;
}

extern "C" int ed_slab_from_rowmajor(ed_hubbard hh, const double *d_x, double *d_xl, void *stream)
{
    const Hubbard *h = reinterpret_cast<const Hubbard *>(hh);
    if (!h || !h->slab || !d_x || !d_xl) return fail(ED_EINVAL, "ed_slab_from_rowmajor: no slab layout or null buffer");
    return slab_from_rowmajor(*h, d_x, d_xl, stream_of(stream));
}

extern "C" int ed_slab_to_rowmajor(ed_hubbard hh, const double *d_xl, double *d_x, void *stream)
{
    const Hubbard *h = reinterpret_cast<const Hubbard *>(hh);
    if (!h || !h->slab || !d_x || !d_xl) return fail(ED_EINVAL, "ed_slab_to_rowmajor: no slab layout or null buffer");
    return slab_to_rowmajor(*h, d_xl, d_x, stream_of(stream));
}

template <int NB, int TR>
static int slab_ceiling_t(const Hubbard &h, int pass, const double *d_xl, hipStream_t st)
{
    const SlabLayout &L = *h.slab;
    const SlabArgs a = slab_args(h, pass == 0);
    const int64_t rows = pass == 0 ? L.du : L.dd, slabs = pass == 0 ? L.nsd : L.nsu;
    const int64_t ntr = (rows + TR - 1) / TR, tiles = slabs * ntr, per = (tiles + 7) / 8;
    if (per * 8 > 0x7fffffffLL) return fail(ED_EUNSUPPORTED, "slab gather ceiling: grid too large");
    // pass 1 gathers over the layout-T scratch, as k_slab_dn does
    const double *src = pass == 0 ? d_xl : L.xt;
    hipLaunchKernelGGL((k_slab_ceiling<NB, TR>), dim3((unsigned)(per * 8)), dim3(8 * TR), 0, st, a, tiles, per, ntr,
                       rows, slabs, pass == 0 ? L.wl : L.wt, src, L.part, L.npart);
    ED_CHECK_LAUNCH();
    return ED_OK;
}

extern "C" int ed_hubbard_slab_gather_ceiling(ed_hubbard hh, int pass, const double *d_xl, void *stream)
{
    const Hubbard *h = reinterpret_cast<const Hubbard *>(hh);
    if (!h || !h->slab || !d_xl || (pass != 0 && pass != 1))
        return fail(ED_EINVAL, "ed_hubbard_slab_gather_ceiling: no slab layout, null buffer or pass not 0/1");
    const SlabLayout &L = *h->slab;
    hipStream_t st = stream_of(stream);
    if (L.tr == 64) return L.nb == 16 ? slab_ceiling_t<16, 64>(*h, pass, d_xl, st) : slab_ceiling_t<8, 64>(*h, pass, d_xl, st);
    return L.nb == 16 ? slab_ceiling_t<16, 32>(*h, pass, d_xl, st) : slab_ceiling_t<8, 32>(*h, pass, d_xl, st);
}

extern "C" int ed_hubbard_slab_set_shape(ed_hubbard hh, int tile_rows, int gathers_in_flight, int tile_order)
{
    Hubbard *h = reinterpret_cast<Hubbard *>(hh);
    if (!h || !h->slab) return fail(ED_EINVAL, "ed_hubbard_slab_set_shape: no slab layout");
    if ((tile_rows != 32 && tile_rows != 64) || (gathers_in_flight != 8 && gathers_in_flight != 16) ||
        (tile_order != 0 && tile_order != 2 && tile_order != 4 && tile_order != 8 && tile_order != 16))
        return fail(ED_EINVAL, "ed_hubbard_slab_set_shape: tile rows 32/64, gathers 8/16, tile order 0/2/4/8/16");
    h->slab->tr = tile_rows;
    h->slab->nb = gathers_in_flight;
    h->slab->swz = tile_order;
    return ED_OK;
}

// Slab autotuner: times the launch shapes of the two passes on the caller's vectors and keeps the
// fastest in the layout.  Stage 1: tile rows (32, 64) x gathers in flight per lane (8, 16) in the
// default tile order; stage 2: the tile order (plain, or runs of 2, 4, 8, 16 adjacent tiles per
// CU slot, slab_tile) with the best stage-1 shape.  Every candidate adds the same terms in the same
// order per element, so y does not depend on the choice.
extern "C" int ed_hubbard_slab_autotune(ed_hubbard hh, const double *d_xl, double *d_yl, int reps,
                                        ed_slab_tuning *best, void *stream)
{
    Hubbard *h = reinterpret_cast<Hubbard *>(hh);
    if (!h || !h->slab || !d_xl || !d_yl || d_xl == d_yl || reps < 1)
        return fail(ED_EINVAL, "ed_hubbard_slab_autotune: no slab layout, null or aliased buffers, or reps < 1");
    SlabLayout &L = *h->slab;
    hipStream_t st = stream_of(stream);
    hipEvent_t e0, e1;
    ED_HIP(hipEventCreate(&e0));
    ED_HIP(hipEventCreate(&e1));
    double cand[8];
    for (double &c : cand) c = -1.0;
    int s = ED_OK;
    auto time_one = [&](float &ms) -> int {
        ED_TRY(slab_apply(*h, d_xl, d_yl, nullptr, d_yl, nullptr, nullptr, 0, 0, nullptr, st, nullptr));
        ED_HIP(hipEventRecord(e0, st));
        for (int r = 0; r < reps; ++r)
            ED_TRY(slab_apply(*h, d_xl, d_yl, nullptr, d_yl, nullptr, nullptr, 0, 0, nullptr, st, nullptr));
        ED_HIP(hipEventRecord(e1, st));
        ED_HIP(hipEventSynchronize(e1));
        ED_HIP(hipEventElapsedTime(&ms, e0, e1));
        ms /= (float)reps;
        return ED_OK;
    };
    float best_ms = -1.0f;
    int btr = L.tr, bnb = L.nb, bswz = L.swz, ci = 0;
    L.swz = 8;
    for (int tr : {32, 64})
        for (int nb : {8, 16}) {
            if (s) break;
            L.tr = tr;
            L.nb = nb;
            float ms = 0.0f;
            if ((s = time_one(ms))) break;
            cand[ci++] = ms;
            if (best_ms < 0.0f || ms < best_ms) {
                best_ms = ms;
                btr = tr;
                bnb = nb;
                bswz = 8;
            }
        }
    for (int swz : {0, 2, 4, 16}) {
        if (s) break;
        L.tr = btr;
        L.nb = bnb;
        L.swz = swz;
        float ms = 0.0f;
        if ((s = time_one(ms))) break;
        cand[ci++] = ms;
        if (ms < best_ms) {
            best_ms = ms;
            bswz = swz;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    L.tr = btr;
    L.nb = bnb;
    L.swz = bswz;
    if (s) return s;
    if (best) {
        best->tile_rows = btr;
        best->gathers_in_flight = bnb;
        best->tile_order = bswz;
        best->ms_per_apply = best_ms;
        for (int i = 0; i < 8; ++i) best->candidate_ms[i] = cand[i];
    }
    return ED_OK;
}

extern "C" int ed_hubbard_apply_slab(ed_hubbard hh, const double *d_xl, double *d_yl, void *stream)
{
    const Hubbard *h = reinterpret_cast<const Hubbard *>(hh);
    if (!h || !h->slab || !d_xl || !d_yl || d_xl == d_yl)
        return fail(ED_EINVAL, "ed_hubbard_apply_slab: no slab layout, null or aliased buffers");
    // the UP partial goes into y, which pass B then overwrites in place
    return slab_apply(*h, d_xl, d_yl, nullptr, d_yl, nullptr, nullptr, 0, 0, nullptr, stream_of(stream), nullptr);
}

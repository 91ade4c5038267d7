// ed_batch.hip -- many independent Lanczos ground-state solves in one stream of launches.
//
// The reference solves one Hubbard instance at a time (Lanczos<T>::eigen,
// MBL_test/include/HubMatrix_GPU.cuh:561-613, called once per sample/sector by the drivers,
// e.g. drude_v2_cal_Kagome_randomHop.cu:129-175).  For sweeps over uneven sectors, twist
// fluxes or disorder realisations that leaves the GPU mostly idle on the small instances and
// pays a host round trip per check on every one.  Here all instances advance together.
//
// Groups.  Instances whose operators have the same structure -- same sector, the same
// H_TBq calls (kind, i, j) in the same order, the same U, dilution mask and per-site U -- and
// differ only in call values (hopping amplitudes, twist phases, on-site potentials) form a
// group of up to 16 copies.  A group's vectors are interleaved, psi[(i_up*dim_dn + i_dn)*G + k]
// for copy k, so one hop descriptor (neighbour rank, sign, call) serves all copies and every
// gather reads G consecutive amplitudes; the call values sit in a per-copy table.  Instances
// that share nothing form groups of one.
//
//   * one launch of each kernel covers every tile of every group (tile lists with per-group
//     offsets, binary search for the owner);
//   * each copy keeps its own Lanczos scalars, active flag and deterministic reductions
//     (per-tile, per-copy partials; one finalize block per copy);
//   * the step number, the active flags and the convergence test live on the device: after
//     every ten steps (the reference's check cadence) a check kernel runs the test of every
//     active copy (Sturm bisection and inverse iteration of its tridiagonal, the same code as
//     the host test of ed_lanczos_eigen, ed_tridiag.hpp) and switches converged copies off, so
//     a graph of several ten-step chunks with their checks is captured once per buffer rotation
//     phase and replayed; the host only reads the count of copies still running between
//     replays.
// Each copy follows the recurrence and convergence test of ed_lanczos_eigen (eigenvalue only),
// with its own instance's start vector (GenRand of the sector length).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "ed_common.hpp"
#include "ed_hubbard.hpp"
#include "ed_hv_otf_body.hpp"
#include "ed_tridiag.hpp"

namespace ed {

constexpr int GMAX = 16;  // copies per group
constexpr int GRMAX = 8;  // DN rows per tile: 4 or 8 (Batch::gr)
constexpr int GUMAX = 4;  // UP rows per tile: 1 or 4 (Batch::gu), one wave per row's calls
constexpr int GCH = 8;    // elements per thread per update chunk

struct GGroup {
    OtfArgs a;               // structure (calls of copy 0), dims, basis, U / mask / per-site U
    const double *vt_up;     // [nops_up][G] x {vf, vb}, TV each (real or complex)
    const double *vt_dn;
    int G, lg, ncopy, cbase; // copies (power of 2), log2 G, real copies, first global copy
    int CW, has_up;          // columns per tile = 256 / G
    int64_t voff;            // group vector offset in doubles
    int64_t nelem;           // dim_up * dim_dn
    int64_t ncb;             // column blocks per row
    int64_t t1_off, t1_n;    // UP tiles (GU rows x CW columns, slab-major)
    int64_t t2_off, t2_n;    // DN tiles (GR rows x CW columns)
    int64_t c_off, c_n;      // update chunks (GCH * CW elements)
    int64_t p_off;           // partials: max(nu_dn, c_n) * G
    int64_t nu_dn;           // DN dot partials per copy: 4-row units x column blocks, any panel shape
    int up_gather, dn_gather; // calls 0..gather-1 gather; the rest are diagonal (the own amplitude)
};



// value of descriptor code (4 * call + {+v_f, -v_f, +v_b, -v_b}) for copy k, from the
// group's value table staged in LDS
template <typename TV>
__device__ __forceinline__ TV gval(const TV *t, int G, uint32_t code, int k)
{
    const TV v = t[((code >> 2) * G + k) * 2 + ((code >> 1) & 1)];
    return (code & 1) ? tv_neg<TV>(v) : v;
}

// copy a group's value table ([nops][G] x {vf, vb}) into LDS
template <typename TV>
__device__ __forceinline__ void stage_table(const double *src, int nops, int G, TV *dst)
{
    const TV *s = reinterpret_cast<const TV *>(src);
    for (int q = threadIdx.x; q < nops * G * 2; q += 256) dst[q] = s[q];
}

// Whether a workgroup's group still has a running copy, the same answer for every thread of the
// workgroup.  The check kernel that runs beside a step launch can switch the group's last copy off
// at any moment, and each wave reading the count itself would read it at its own time: one wave
// could leave while another stays on and reads the row descriptors and the count that the leaver
// should have written to LDS, gathering from wherever those stale words point.  Thread 0 reads
// the count once and the workgroup leaves or stays as a whole.
__device__ __forceinline__ bool group_running(const int *gact, int grp, int *s_flag)
{
    if (threadIdx.x == 0) *s_flag = gact[grp];
    __syncthreads();
    return *s_flag != 0;
}

// descriptor of call op in the row/column of state s with rank self: (col << 8 | code), or ~0
__device__ __forceinline__ uint32_t gdesc(const OtfOp &op, int o, uint32_t s, uint32_t self,
                                          const int32_t *__restrict__ addr)
{
    const uint32_t oi = (s >> op.i) & 1u, oj = (s >> op.j) & 1u;
    if (op.kind == ED_OP_DENSITY) return oi ? ((self << 8) | (uint32_t)(4 * o)) : 0xffffffffu;
    if (op.kind == ED_OP_DENSITY2) return (oi & oj) ? ((self << 8) | (uint32_t)(4 * o)) : 0xffffffffu;
    if (oi == oj) return 0xffffffffu;
    const uint32_t code = (oi ? 0u : 2u) | (uint32_t)(__popc(s & op.between) & 1);
    return ((uint32_t)addr[s ^ op.m] << 8) | (uint32_t)(4 * o) | code;
}

// Pass 1: UP term of GU rows x CW columns x G copies -> y (the partial), x scaled by 1/b_N.
// Tiles run slab-major (all rows of a CW-column slab, then the next slab), so the rows the
// hops gather from -- dim_up x CW x G amplitudes, 3.7 MB for the 924 x 924 x 16 sector --
// stay in the XCD's L2 while the slab is processed.
template <typename V, typename TV, int GU, bool DIAG>
__global__ __launch_bounds__(256) void k_g_up(const GGroup *__restrict__ gr, const int *__restrict__ tgrp, int64_t ntiles, int64_t per_xcd,
                                              const int *__restrict__ active, const int *__restrict__ gact,
                                              const double *__restrict__ bdev, int ldb,
                                              const int *__restrict__ dstep, const double *x, double *y, int kp)
{
    extern __shared__ double s_upv[]; // value table, TV [nops_up][G][2]
    __shared__ uint32_t s_desc[GUMAX][64];
    __shared__ int s_cnt[GUMAX], s_nh[GUMAX];
    __shared__ int s_run;
    const int64_t b = blockIdx.x;
    const int64_t T = (b & 7) * per_xcd + (b >> 3);
    if (T >= ntiles) return; // past the tiles (uniform)
    if (!group_running(gact, tgrp[T], &s_run)) return; // every copy of the group is done
    const GGroup g = gr[tgrp[T]];
    if (!g.has_up) return;
    TV *s_vt = reinterpret_cast<TV *>(s_upv);
    stage_table<TV>(g.vt_up, g.a.up.nops, g.G, s_vt);
    const int64_t du = g.a.dim_up, dd = g.a.dim_dn;
    const int64_t nrg = (du + (int64_t)GU * kp - 1) / ((int64_t)GU * kp);
    const int64_t tl = T - g.t1_off, cb = tl / nrg, rg = tl - cb * nrg;
    const int k = threadIdx.x & (g.G - 1), c = threadIdx.x >> g.lg;
    const int64_t d = cb * g.CW + c;
    const int gi = g.cbase + k;
    const bool live = k < g.ncopy && d < dd && active[gi];
    const double sc = live ? 1.0 / bdev[(int64_t)*dstep * ldb + gi] : 0.0;
    const V *xv = reinterpret_cast<const V *>(x + g.voff);
    V *yv = reinterpret_cast<V *>(y + g.voff);
    const uint32_t ddu = (uint32_t)dd, lg = (uint32_t)g.lg, ku = (uint32_t)k, du32 = (uint32_t)d;
    // kp panels of GU rows; the value table is staged once per workgroup
    for (int pn = 0; pn < kp; ++pn) {
        const int64_t u0 = (rg * kp + pn) * GU;
        if (u0 >= du) break; // uniform
        if (pn > 0) __syncthreads(); // the previous panel's hop lists are no longer read
        if ((threadIdx.x >> 6) < GU) { // wave r evaluates the calls of row u0 + r
            const int r = threadIdx.x >> 6, lane = threadIdx.x & 63;
            const int64_t u = u0 + r;
            uint32_t dsc = 0xffffffffu;
            if (u < du && lane < g.a.up.nops)
                dsc = gdesc(g.a.up.ops[lane], lane, (uint32_t)g.a.state_up[u], (uint32_t)u, g.a.up.addr);
            const bool has = dsc != 0xffffffffu;
            const uint64_t bal = __ballot(has);
            if (has) s_desc[r][__popcll(bal & ((1ull << lane) - 1ull))] = dsc;
            if (lane == 0) {
                s_cnt[r] = __popcll(bal);
                // entries of the gathering calls; the rest (trailing diagonal calls) point at the row
                if (DIAG) s_nh[r] = g.up_gather >= 64 ? __popcll(bal) : __popcll(bal & ((1ull << g.up_gather) - 1ull));
            }
        }
        __syncthreads();
        if (!live) continue;
        for (int r = 0; r < GU; ++r) {
            const int64_t u = u0 + r;
            if (u >= du) break;
            const int n = s_cnt[r], nh = DIAG ? s_nh[r] : n;
            V acc = VO<V>::zero();
            for (int j0 = 0; j0 < nh; j0 += 8) {
                V xl[8];
                uint32_t ds[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    ds[q] = (j0 + q < nh) ? s_desc[r][j0 + q] : 0u;
                    if (j0 + q < nh) xl[q] = xv[(((ds[q] >> 8) * ddu + du32) << lg) + ku];
                }
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    if (j0 + q < nh)
                        acc = omac(gval<TV>(s_vt, g.G, ds[q] & 0xffu, k), VO<V>::scale(sc, xl[q]), acc);
            }
            // the trailing diagonal calls' entries (Make_N, Insert_2Nop after the last hop) read
            // the row's own amplitude: one load for all of them (the same value, the same sum)
            if (DIAG && n > nh) {
                const V xs = VO<V>::scale(sc, xv[(((uint32_t)u * ddu + du32) << lg) + ku]);
                for (int j = nh; j < n; ++j) acc = omac(gval<TV>(s_vt, g.G, s_desc[r][j] & 0xffu, k), xs, acc);
            }
            yv[(((uint32_t)u * ddu + du32) << lg) + ku] = acc;
        }
    }
}

// Pass 2 of one step on a tile of GR rows x CW columns x G copies (row-panel order, so the
// rows the DN hops gather from stay in L2): y = (UP partial in y) + DN term + diagonal
// - b_N v_{N-1}, with v = x / b_N and v_{N-1} = xp / b_{N-1} recomputed from the previous
// start vector instead of being stored; per-copy partials of Re<v, y>.  Per element the UP
// calls, then the DN calls, then the diagonal, as in the single-operator kernels.  Step 0 has
// no v_{N-1} term.
template <typename V, typename TV, bool AXPY, int GR>
__global__ __launch_bounds__(256) void k_g_dn(const GGroup *__restrict__ gr, const int *__restrict__ tgrp, int64_t ntiles, int64_t per_xcd,
                                              const int *__restrict__ active, const int *__restrict__ gact,
                                              const double *__restrict__ bdev, int ldb,
                                              const int *__restrict__ dstep, const double *x, const double *xp,
                                              double *y, double *partial, int kp)
{
    extern __shared__ double s_dnv[]; // value table TV [nops_dn][G][2], then descriptors [nops_dn][CW]
    __shared__ double s_red[256];
    __shared__ int s_run;
    const int64_t b = blockIdx.x;
    const int64_t T = (b & 7) * per_xcd + (b >> 3);
    if (T >= ntiles) return; // past the tiles (uniform)
    if (!group_running(gact, tgrp[T], &s_run)) return; // every copy of the group is done
    const GGroup g = gr[tgrp[T]];
    TV *s_vt = reinterpret_cast<TV *>(s_dnv);
    uint32_t *s_gd = reinterpret_cast<uint32_t *>(s_vt + g.a.dn.nops * g.G * 2);
    stage_table<TV>(g.vt_dn, g.a.dn.nops, g.G, s_vt);
    const int64_t tl = T - g.t2_off, rp = tl / g.ncb, cb = tl - rp * g.ncb;
    const int64_t du = g.a.dim_up, dd = g.a.dim_dn;
    const int k = threadIdx.x & (g.G - 1), c = threadIdx.x >> g.lg;
    const int64_t d = cb * g.CW + c;
    const int nops = g.a.dn.nops;
    const uint32_t sd = d < dd ? (uint32_t)g.a.state_dn[d] : 0u;
    for (int o = k; o < nops; o += g.G) // the G lanes of a column share its descriptors
        s_gd[o * g.CW + c] = d < dd ? gdesc(g.a.dn.ops[o], o, sd, (uint32_t)d, g.a.dn.addr) : 0xffffffffu;
    __syncthreads();
    const int gi = g.cbase + k;
    const bool live = k < g.ncopy && d < dd && active[gi];
    // kp panels of GR rows share the column descriptors and the value table; every thread walks
    // the panels (the per-unit sums below take barriers), threads without an element add nothing
    for (int pn = 0; pn < kp; ++pn) {
        const int64_t u0 = (rp * kp + pn) * GR;
        if (u0 >= du) break; // uniform over the workgroup
        double dotu[GR / 4];
#pragma unroll
        for (int q = 0; q < GR / 4; ++q) dotu[q] = 0.0;
        if (live) {
            const int step = *dstep;
            const double bN = bdev[(int64_t)step * ldb + gi], sc = 1.0 / bN;
            const double scp = AXPY ? 1.0 / bdev[(int64_t)(step - 1) * ldb + gi] : 0.0;
            const V *xv = reinterpret_cast<const V *>(x + g.voff);
            const V *xpv = reinterpret_cast<const V *>(xp + g.voff);
            V *yy = reinterpret_cast<V *>(y + g.voff);
            // 32-bit element offsets inside the group (ed_batch_create checks nelem * G < 2^31):
            // row r's amplitude of column col and copy k sits at ((rowbase[r] + col) << lg) + k
            const uint32_t ddu = (uint32_t)dd, lg = (uint32_t)g.lg, ku = (uint32_t)k;
            uint32_t rowbase[GR];
#pragma unroll
            for (int r = 0; r < GR; ++r) rowbase[r] = (uint32_t)((u0 + r < du) ? u0 + r : u0) * ddu;
            V acc[GR];
#pragma unroll
            for (int r = 0; r < GR; ++r)
                acc[r] = (g.has_up && u0 + r < du) ? yy[((rowbase[r] + (uint32_t)d) << lg) + ku] : VO<V>::zero();
            // no branches around the gathers, so the KB x GR loads of a chunk are all in flight: a
            // column without an element of a call loads its own amplitude (contiguous with its
            // neighbours' and needed below anyway) and adds 0 times it, which leaves the sum
            // unchanged bit for bit (finite x; the sums start at +0 or the pass-1 value, never -0)
            constexpr int KB = 4;
            const TV tzero = tv_make<TV>(0.0, 0.0);
            const int nhop = g.dn_gather; // calls from nhop on are diagonal: no gathers, below
            for (int o0 = 0; o0 < nhop; o0 += KB) {
                uint32_t col[KB];
                TV av[KB];
                V xl[KB][GR];
#pragma unroll
                for (int q = 0; q < KB; ++q) {
                    const uint32_t e = (o0 + q < nhop) ? s_gd[(o0 + q) * g.CW + c] : 0xffffffffu;
                    const bool on = e != 0xffffffffu;
                    col[q] = on ? (e >> 8) : (uint32_t)d;
                    const TV a = gval<TV>(s_vt, g.G, on ? (e & 0xffu) : 0u, k);
                    av[q] = on ? a : tzero;
                }
#pragma unroll
                for (int q = 0; q < KB; ++q)
#pragma unroll
                    for (int r = 0; r < GR; ++r) xl[q][r] = xv[((rowbase[r] + col[q]) << lg) + ku];
#pragma unroll
                for (int q = 0; q < KB; ++q)
#pragma unroll
                    for (int r = 0; r < GR; ++r) acc[r] = omac(av[q], VO<V>::scale(sc, xl[q][r]), acc[r]);
            }
            // the trailing diagonal calls (Make_N, Insert_2Nop after the last hop) all read the
            // row's own amplitude: loaded once per row here instead of once per call (the same
            // value; a column without the call's element adds 0 times it as above: the same sums)
            if (nhop < nops) {
#pragma unroll
                for (int r = 0; r < GR; ++r) {
                    const V xr = VO<V>::scale(sc, xv[((rowbase[r] + (uint32_t)d) << lg) + ku]);
                    for (int o = nhop; o < nops; ++o) {
                        const uint32_t e = s_gd[o * g.CW + c];
                        const bool on = e != 0xffffffffu;
                        const TV a = gval<TV>(s_vt, g.G, on ? (e & 0xffu) : 0u, k);
                        acc[r] = omac(on ? a : tzero, xr, acc[r]);
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < GR; ++r) {
                const int64_t u = u0 + r;
                if (u >= du) continue;
                const uint32_t i = ((rowbase[r] + (uint32_t)d) << lg) + ku;
                const V xd = VO<V>::scale(sc, xv[i]);
                if (g.a.has_diag) acc[r] = VO<V>::fmr(diag_of(g.a.state_up[u], (int32_t)sd, g.a), xd, acc[r]);
                V o = acc[r];
                if (AXPY) o = VO<V>::fmr(-bN, VO<V>::scale(scp, xpv[i]), o);
                yy[i] = o;
                dotu[r / 4] += VO<V>::dotre(xd, o);
            }
        }
        // per-copy sums of each 4-row unit in a fixed order, thread (c, k) -> s_red[c * G + k], into
        // partial slot (unit * ncb + cb): the same slots and values whatever GR and kp are
#pragma unroll
        for (int q = 0; q < GR / 4; ++q) {
            const int64_t unit = u0 / 4 + q;
            if (unit * 4 >= du) break; // uniform
            s_red[threadIdx.x] = dotu[q];
            __syncthreads();
            if (threadIdx.x < g.G) {
                double sm = 0.0;
                for (int cc = 0; cc < g.CW; ++cc) sm += s_red[cc * g.G + threadIdx.x];
                partial[g.p_off + (unit * g.ncb + cb) * g.G + threadIdx.x] = sm;
            }
            __syncthreads();
        }
    }
}

// f = -a_k v + f with v = x / b_N over GCH * CW elements x G copies; per-copy partials of |f|^2
__global__ __launch_bounds__(256) void k_g_update(const GGroup *__restrict__ gr, const int *__restrict__ tgrp, int64_t ntiles,
                                                  const int *__restrict__ active, const int *__restrict__ gact,
                                                  const double *__restrict__ adev, const double *__restrict__ bdev,
                                                  int lda, const int *__restrict__ dstep, int cvec, const double *x,
                                                  double *f, double *partial, const double *__restrict__ zdev,
                                                  const int *__restrict__ lim, double *acc)
{
    __shared__ double s_red[256];
    __shared__ int s_run;
    const int64_t T = blockIdx.x;
    if (T >= ntiles) return; // past the tiles (uniform)
    if (!group_running(gact, tgrp[T], &s_run)) return; // every copy of the group is done
    const GGroup g = gr[tgrp[T]];
    const int64_t ch = T - g.c_off;
    const int k = threadIdx.x & (g.G - 1), c = threadIdx.x >> g.lg;
    const int gi = g.cbase + k;
    double s = 0.0;
    if (k < g.ncopy && active[gi]) {
        const int N = *dstep;
        const double a = -adev[(int64_t)N * lda + gi];
        const double sc = 1.0 / bdev[(int64_t)N * lda + gi];
        const int NC = cvec ? 2 : 1;
        // the rebuild run: evec += z_N v_N over the copy's first lim steps
        const bool ritz = acc != nullptr && N < lim[gi];
        const double zc = ritz ? zdev[(int64_t)N * lda + gi] : 0.0;
#pragma unroll
        for (int j = 0; j < GCH; ++j) {
            const int64_t e = (ch * GCH + j) * g.CW + c;
            if (e < g.nelem) {
                const int64_t i = g.voff + (e * g.G + k) * NC;
                for (int q = 0; q < NC; ++q) {
                    const double v = sc * x[i + q];
                    const double r = fma(a, v, f[i + q]);
                    f[i + q] = r;
                    s = fma(r, r, s);
                    if (ritz) acc[i + q] = fma(zc, v, acc[i + q]);
                }
            }
        }
    }
    s_red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < g.G) {
        double t = 0.0;
        for (int cc = 0; cc < g.CW; ++cc) t += s_red[cc * g.G + threadIdx.x];
        partial[g.p_off + ch * g.G + threadIdx.x] = t;
    }
}

// one block per global copy: dst[step(+1)*ld + copy] = sum (which 0) or sqrt(sum) (which 1);
// a and b are stored step-major, so the entries added since a check are one contiguous range
__global__ __launch_bounds__(1024) void k_g_finalize(const GGroup *__restrict__ gr, const int *__restrict__ cgroup,
                                                     const int *__restrict__ active, const double *__restrict__ partial,
                                                     int which, double *dst, int ld, const int *dstep)
{
    __shared__ double red[16];
    __shared__ int s_run;
    const int gi = blockIdx.x;
    // one read of the copy's flag for the whole workgroup (a check beside this launch may clear it)
    if (!group_running(active, gi, &s_run)) return;
    const GGroup &g = gr[cgroup[gi]];
    const int k = gi - g.cbase;
    const int64_t np = which == 0 ? g.nu_dn : g.c_n;
    double s = 0.0;
    for (int64_t t = threadIdx.x; t < np; t += 1024) s += partial[g.p_off + t * g.G + k];
    s = block_sum<1024>(s, red);
    if (threadIdx.x == 0) {
        const int step = *dstep;
        if (which == 0) dst[(int64_t)step * ld + gi] = s;
        else dst[(int64_t)(step + 1) * ld + gi] = sqrt(s);
    }
}

__global__ void k_g_bump(int *dstep) { dstep[0] += 1; }

// start vectors: copies 0..ncopy-1 of a group <- src (their sector's GenRand vector); NC
// doubles per element
__global__ __launch_bounds__(256) void k_g_scatter(int64_t nelem, int G, int ncopy, int NC,
                                                   const double *__restrict__ src, double *__restrict__ dst)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= nelem) return;
    for (int k = 0; k < ncopy; ++k)
        for (int q = 0; q < NC; ++q) dst[(e * G + k) * NC + q] = src[e * NC + q];
}

// b_0 row (step 0 of the step-major b array) of one group: |f| for its copies, 1 for padding
__global__ void k_g_b0(int cbase, int ncopy, int G, const double *__restrict__ nrm, double *__restrict__ bdev)
{
    const int k = threadIdx.x;
    if (k >= G) return;
    const double v = *nrm;
    bdev[cbase + k] = (k < ncopy && v > 0.0) ? v : 1.0;
}

// The rebuild run's stand-in for the convergence test: after a chunk, copies that have run their
// lim steps are switched off
__global__ void k_g_limit(int nc, const int *__restrict__ dchk, int chunk, int *active, const int *__restrict__ cgroup,
                          int *gact, const int *__restrict__ lim, int *remaining)
{
    const int c = blockIdx.x * 64 + threadIdx.x;
    if (c >= nc || !active[c]) return;
    if (*dchk + chunk >= lim[c]) {
        active[c] = 0;
        atomicSub(&gact[cgroup[c]], 1);
        atomicSub(remaining, 1);
    }
}

// copy k of a group's interleaved vector -> dst (one instance, NC doubles per element)
__global__ __launch_bounds__(256) void k_g_gather(int64_t nelem, int G, int k, int NC, const double *__restrict__ src,
                                                  double *__restrict__ dst)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= nelem) return;
    for (int q = 0; q < NC; ++q) dst[e * NC + q] = src[(e * G + k) * NC + q];
}

struct Batch {
    int n = 0, ng = 0;
    bool cvec = false, ctab = false;
    std::vector<GGroup> host;
    std::vector<int> copy_of;  // instance -> global copy index
    std::vector<int> inst_of;  // global copy index -> instance (-1: padding)
    std::vector<int> cgroup;   // global copy -> group
    int ncopies = 0;
    GGroup *dgr = nullptr;
    int *dcgroup = nullptr;
    int *dtg1 = nullptr, *dtg2 = nullptr, *dtgc = nullptr; // tile -> group of the UP, DN, update grids
    double *dvt = nullptr;
    int64_t total_real = 0, t1 = 0, t2 = 0, tc = 0, pcap = 0;
    int maxnops_dn = 1;
    size_t lds_up = 0, lds_dn = 0; // dynamic LDS of the UP / DN kernels (largest group)
    int gr = 4, gu = 1;            // DN / UP rows per panel (4 or 8 / 1 or 4)
    bool up_diag = false;          // some group's UP calls end in diagonal calls (k_g_up<..., DIAG>)
    int kp = 4, kpu = 4;           // DN / UP panels per workgroup (ed_batch_autotune, ed_batch_set_shape)
    ~Batch()
    {
        dfree(dgr);
        dfree(dcgroup);
        dfree(dtg1);
        dfree(dtg2);
        dfree(dtgc);
        dfree(dvt);
    }
};

static bool same_structure(const Hubbard *a, const Hubbard *b, const std::vector<OtfOp> &ua,
                           const std::vector<OtfOp> &da, const std::vector<OtfOp> &ub, const std::vector<OtfOp> &db,
                           const std::vector<double> &uva, const std::vector<double> &uvb)
{
    if (a->N != b->N || a->otf.dim_up != b->otf.dim_up || a->otf.dim_dn != b->otf.dim_dn ||
        a->otf.state_up != b->otf.state_up || a->otf.state_dn != b->otf.state_dn || a->U != b->U ||
        a->mask != b->mask || ua.size() != ub.size() || da.size() != db.size() || uva != uvb ||
        a->otf.has_diag != b->otf.has_diag)
        return false;
    auto eq = [](const std::vector<OtfOp> &x, const std::vector<OtfOp> &y) {
        for (size_t o = 0; o < x.size(); ++o)
            if (x[o].kind != y[o].kind || x[o].i != y[o].i || x[o].j != y[o].j) return false;
        return true;
    };
    return eq(ua, ub) && eq(da, db);
}

// Tile tables of the UP and DN grids for the current panel shape (gu x kpu UP rows, gr x kp DN rows
// per workgroup): each group's tile range, the tile -> group maps and the partial-sum slots; then
// the group descriptors are uploaded.  Rebuilt when the shape changes.
static int batch_tiles(Batch *B)
{
    int64_t t1 = 0, t2 = 0, poff = 0;
    for (GGroup &g : B->host) {
        const int64_t du = g.a.dim_up;
        g.t1_off = t1;
        g.t1_n = ((du + (int64_t)B->gu * B->kpu - 1) / ((int64_t)B->gu * B->kpu)) * g.ncb;
        t1 += g.t1_n;
        g.t2_off = t2;
        g.t2_n = ((du + (int64_t)B->gr * B->kp - 1) / ((int64_t)B->gr * B->kp)) * g.ncb;
        t2 += g.t2_n;
        g.nu_dn = ((du + 3) / 4) * g.ncb;
        g.p_off = poff;
        poff += std::max(g.nu_dn, g.c_n) * g.G;
    }
    B->t1 = t1;
    B->t2 = t2;
    B->pcap = poff;
    std::vector<int> tg1((size_t)t1), tg2((size_t)t2);
    for (int gq = 0; gq < B->ng; ++gq) {
        const GGroup &g = B->host[gq];
        std::fill(tg1.begin() + g.t1_off, tg1.begin() + g.t1_off + g.t1_n, gq);
        std::fill(tg2.begin() + g.t2_off, tg2.begin() + g.t2_off + g.t2_n, gq);
    }
    dfree(B->dtg1);
    dfree(B->dtg2);
    B->dtg1 = B->dtg2 = nullptr;
    ED_TRY(dmalloc((void **)&B->dtg1, sizeof(int) * std::max<int64_t>(t1, 1)));
    ED_TRY(dmalloc((void **)&B->dtg2, sizeof(int) * std::max<int64_t>(t2, 1)));
    if (hipMemcpy(B->dtg1, tg1.data(), sizeof(int) * t1, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(B->dtg2, tg2.data(), sizeof(int) * t2, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(B->dgr, B->host.data(), sizeof(GGroup) * B->ng, hipMemcpyHostToDevice) != hipSuccess)
        return fail(ED_EHIP, "batch: tile table upload failed");
    return ED_OK;
}

} // namespace ed

using namespace ed;

extern "C" int ed_batch_create(int ninst, const ed_hubbard *ops, int complex_vectors, ed_batch *out)
{
    if (!out || ninst < 1 || !ops) return fail(ED_EINVAL, "ed_batch_create: bad arguments");
    *out = nullptr;
    Batch *B = new Batch();
    B->n = ninst;
    B->cvec = complex_vectors != 0;
    std::vector<const Hubbard *> hs(ninst);
    std::vector<std::vector<OtfOp>> opu(ninst), opd(ninst);
    std::vector<std::vector<double>> uv(ninst);
    for (int i = 0; i < ninst; ++i) {
        const Hubbard *h = reinterpret_cast<const Hubbard *>(ops[i]);
        if (!h || h->algo != 3 || !h->otf.up.ops_own || !h->otf.dn.ops_own) {
            delete B;
            return fail(ED_EINVAL, "ed_batch_create: instance %d must come from ed_hubbard_create_ops", i);
        }
        if (i == 0) B->ctab = h->cplx_tables;
        if (h->cplx_tables != B->ctab || (h->cplx_tables && !B->cvec)) {
            delete B;
            return fail(ED_EINVAL, "ed_batch_create: instances must share the table type (complex needs complex vectors)");
        }
        if (h->otf.dim_dn >= (1 << 24) || h->otf.dim_up >= (1 << 24)) {
            delete B;
            return fail(ED_EUNSUPPORTED, "ed_batch_create: sector dimension %lld x %lld too large for the batch",
                        (long long)h->otf.dim_up, (long long)h->otf.dim_dn);
        }
        hs[i] = h;
        // host copies kept by ed_hubbard_create_ops (no device round trip per instance)
        if ((int)h->h_ops_up.size() != h->otf.up.nops || (int)h->h_ops_dn.size() != h->otf.dn.nops ||
            (h->d_uvec && (int)h->h_uvec.size() != h->N)) {
            delete B;
            return fail(ED_EINVAL, "ed_batch_create: instance %d has no host copy of its operator", i);
        }
        opu[i] = h->h_ops_up;
        opd[i] = h->h_ops_dn;
        if (h->d_uvec) uv[i] = h->h_uvec;
    }
    // greedy grouping, first-seen order
    std::vector<std::vector<int>> groups;
    for (int i = 0; i < ninst; ++i) {
        int gsel = -1;
        for (int gq = 0; gq < (int)groups.size() && gsel < 0; ++gq) {
            const int r = groups[gq][0];
            if ((int)groups[gq].size() < GMAX && same_structure(hs[r], hs[i], opu[r], opd[r], opu[i], opd[i], uv[r], uv[i]))
                gsel = gq;
        }
        if (gsel < 0) {
            groups.push_back({});
            gsel = (int)groups.size() - 1;
        }
        groups[gsel].push_back(i);
    }
    const size_t tv = B->ctab ? 2 : 1; // doubles per value
    std::vector<double> vt;
    std::vector<int64_t> vt_up_off, vt_dn_off;
    int64_t voff = 0, tc = 0;
    B->copy_of.assign(ninst, -1);
    for (const auto &mem : groups) {
        const Hubbard *h = hs[mem[0]];
        GGroup g{};
        g.a = h->otf;
        g.ncopy = (int)mem.size();
        g.G = 1;
        g.lg = 0;
        // DN descriptors of CW = 256 / G columns and the value table must fit 62 KB of LDS
        auto dn_lds = [&](int G) {
            return (size_t)h->otf.dn.nops * (256 / G) * 4 + (size_t)h->otf.dn.nops * G * 2 * 8 * tv;
        };
        while (g.G < g.ncopy || dn_lds(g.G) > 63488) g.G <<= 1, ++g.lg;
        B->lds_dn = std::max(B->lds_dn, dn_lds(g.G));
        B->lds_up = std::max(B->lds_up, (size_t)h->otf.up.nops * g.G * 2 * 8 * tv);
        g.CW = 256 / g.G;
        g.cbase = B->ncopies;
        for (int k = 0; k < g.G; ++k) {
            B->inst_of.push_back(k < g.ncopy ? mem[k] : -1);
            B->cgroup.push_back((int)B->host.size());
            if (k < g.ncopy) B->copy_of[mem[k]] = g.cbase + k;
        }
        B->ncopies += g.G;
        g.has_up = h->otf.up.nops > 0 ? 1 : 0;
        const int64_t du = h->otf.dim_up, dd = h->otf.dim_dn;
        g.nelem = du * dd;
        if (g.nelem * g.G >= ((int64_t)1 << 31)) {
            delete B;
            return fail(ED_EUNSUPPORTED, "ed_batch_create: a group of %d copies of a %lld-element sector is too large for "
                        "the batch (use ed_lanczos_eigen per instance)", g.G, (long long)g.nelem);
        }
        const int64_t nreal = g.nelem * g.G * (B->cvec ? 2 : 1);
        g.voff = voff;
        voff += (nreal + 31) / 32 * 32;
        g.ncb = (dd + g.CW - 1) / g.CW;
        g.c_off = tc;
        g.c_n = (g.nelem + (int64_t)GCH * g.CW - 1) / ((int64_t)GCH * g.CW);
        tc += g.c_n;
        B->maxnops_dn = std::max(B->maxnops_dn, h->otf.dn.nops);
        // calls from up_gather / dn_gather on are all diagonal (Make_N / Insert_2Nop after the
        // last hop): the kernels take the row's own amplitude once for them
        auto gather_end = [](const OtfOp *ops, int nops) {
            int e = nops;
            while (e > 0 && (ops[e - 1].kind == ED_OP_DENSITY || ops[e - 1].kind == ED_OP_DENSITY2)) --e;
            return e;
        };
        g.up_gather = gather_end(opu[mem[0]].data(), h->otf.up.nops);
        g.dn_gather = gather_end(opd[mem[0]].data(), h->otf.dn.nops);
        if (g.up_gather < h->otf.up.nops) B->up_diag = true;
        // value tables: [call][copy] x {vf, vb}
        auto table = [&](bool up) {
            const int nops = up ? h->otf.up.nops : h->otf.dn.nops;
            const int64_t off = (int64_t)vt.size();
            vt.resize(vt.size() + (size_t)nops * g.G * 2 * tv, 0.0);
            for (int o = 0; o < nops; ++o)
                for (int k = 0; k < g.ncopy; ++k) {
                    const OtfOp &op = up ? opu[mem[k]][o] : opd[mem[k]][o];
                    double *p = &vt[off + ((size_t)o * g.G + k) * 2 * tv];
                    p[0] = op.vf_re;
                    if (tv == 2) p[1] = op.vf_im;
                    p[tv] = op.vb_re;
                    if (tv == 2) p[tv + 1] = op.vb_im;
                }
            return off;
        };
        vt_up_off.push_back(table(true));
        vt_dn_off.push_back(table(false));
        B->host.push_back(g);
    }
    B->ng = (int)B->host.size();
    B->total_real = voff;
    B->tc = tc;
    int s = dmalloc((void **)&B->dvt, sizeof(double) * std::max<size_t>(vt.size(), 1));
    if (!s) s = dmalloc((void **)&B->dgr, sizeof(GGroup) * B->ng);
    if (!s) s = dmalloc((void **)&B->dcgroup, sizeof(int) * B->ncopies);
    std::vector<int> tgc((size_t)B->tc);
    for (int gq = 0; gq < B->ng; ++gq) {
        const GGroup &g = B->host[gq];
        std::fill(tgc.begin() + g.c_off, tgc.begin() + g.c_off + g.c_n, gq);
    }
    if (!s) s = dmalloc((void **)&B->dtgc, sizeof(int) * std::max<int64_t>(B->tc, 1));
    if (s) {
        delete B;
        return s;
    }
    if (hipMemcpy(B->dtgc, tgc.data(), sizeof(int) * B->tc, hipMemcpyHostToDevice) != hipSuccess) {
        delete B;
        return fail(ED_EHIP, "ed_batch_create: tile table upload failed");
    }
    for (int gq = 0; gq < B->ng; ++gq) {
        B->host[gq].vt_up = B->dvt + vt_up_off[gq];
        B->host[gq].vt_dn = B->dvt + vt_dn_off[gq];
    }
    if ((!vt.empty() && hipMemcpy(B->dvt, vt.data(), sizeof(double) * vt.size(), hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemcpy(B->dcgroup, B->cgroup.data(), sizeof(int) * B->ncopies, hipMemcpyHostToDevice) != hipSuccess) {
        delete B;
        return fail(ED_EHIP, "ed_batch_create: upload failed");
    }
    if ((s = batch_tiles(B))) { // also uploads the group descriptors
        delete B;
        return s;
    }
    *out = reinterpret_cast<ed_batch>(B);
    return ED_OK;
}

extern "C" int ed_batch_destroy(ed_batch bb)
{
    delete reinterpret_cast<Batch *>(bb);
    return ED_OK;
}

// convergence state of one copy (device)
struct CopyCheck {
    double anorm, ev;
    int processed, last, conv, pad;
};

// The convergence test of every active copy after a chunk (ed_lanczos_eigen's test, HubMatrix_GPU.cuh:
// 588-597): the copy's a[0..N-1], b[0..N] gathered from the step-major rows, anorm updated over the
// new steps, the lowest eigenvalue of T by a 64-point multisection of the Gershgorin interval (each
// round keeps the pair of neighbouring points where the Sturm count turns positive, until the two
// ends are adjacent doubles: the bracket the host's bisection ends on, ed_tridiag.hpp), then its
// vector by the host's inverse iteration; a copy that passes (or meets an invariant subspace, or
// has run max_steps) is switched off, and its group's count of running copies drops.  One
// wavefront per copy; the copy's a, b and the inverse iteration's five work arrays (the
// eigenvector overwrites the last) and flags sit in LDS (LDS: 7 ldz doubles + ldz bytes) or in
// ws / wsf (the same per copy).
template <bool LDS>
__global__ __launch_bounds__(64) void k_g_check(int nc, int ld, const double *__restrict__ adev,
                                                const double *__restrict__ bdev, const int *__restrict__ dchk,
                                                int chunk, int *active, const int *__restrict__ cgroup, int *gact,
                                                double tol, int max_steps, CopyCheck *cc, double *ws, char *wsf,
                                                int ldz, int *remaining)
{
#pragma clang fp contract(off)
    extern __shared__ double s_chk[];
    const int c = blockIdx.x, lane = threadIdx.x;
    if (c >= nc || !active[c]) return; // uniform over the wavefront
    const int Ne = *dchk + chunk, Nl = Ne - 1; // steps done at the end of the chunk checked
    double *a = LDS ? s_chk : ws + (size_t)c * 7 * ldz, *b = a + ldz, *w5 = b + ldz;
    char *sw = LDS ? reinterpret_cast<char *>(s_chk + 7 * (size_t)ldz) : wsf + (size_t)c * ldz;
    for (int k = lane; k <= Nl; k += 64) a[k] = adev[(int64_t)k * ld + c];
    for (int k = lane; k <= Ne; k += 64) b[k] = bdev[(int64_t)k * ld + c];
    __syncthreads();
    CopyCheck st = cc[c];
    int M = Nl;
    for (int k = st.processed; k <= Nl; ++k)
        if (!(b[k + 1] > 0.0) || !isfinite(a[k])) {
            M = k;
            break;
        }
    for (int k = st.processed; k <= M; ++k) {
        if (k == 0) st.anorm = fabs(a[0] + b[1]);
        else st.anorm = fmax(st.anorm, b[k] + fabs(a[k] + b[k + 1]));
    }
    st.processed = M + 1;
    const int n = M + 1;
    const double *d = a, *e = b + 1;
    double lo = d[0], hi = d[0];
    for (int i = 0; i < n; ++i) {
        const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < n - 1 ? fabs(e[i]) : 0.0);
        lo = fmin(lo, d[i] - r);
        hi = fmax(hi, d[i] + r);
    }
    const double pad = 1e-12 * (hi - lo + 1.0);
    lo -= pad;
    hi += pad;
    for (int round = 0; round < 64; ++round) {
        const double mid = 0.5 * (lo + hi);
        if (!(mid > lo && mid < hi)) break; // adjacent doubles (uniform)
        const double w = hi - lo;
        const double xj = lo + w * ((double)(lane + 1) / 65.0);
        const unsigned long long above = __ballot(tri_sturm_below(n, d, e, xj) > 0);
        if (above == 0ull) {
            lo = lo + w * (64.0 / 65.0);
        } else {
            const int j = __ffsll((long long)above) - 1;
            const double xh = lo + w * ((double)(j + 1) / 65.0), xl = j > 0 ? lo + w * ((double)j / 65.0) : lo;
            lo = xl;
            hi = xh;
        }
    }
    if (lane != 0) return;
    const double lam = 0.5 * (lo + hi);
    double *z = w5 + 4 * (size_t)n; // the iteration's own vector, overwritten in place
    tri_inverse_iteration(n, d, e, lam, z, w5, sw);
    st.ev = lam;
    st.last = M;
    bool stop = false, conv = false;
    if (M < Nl || !(b[M + 1] > 0.0)) stop = conv = true;
    else if (fabs(b[M + 1] * z[M]) < tol * st.anorm && fabs(z[0]) > tol * st.anorm) stop = conv = true;
    else if (Ne >= max_steps) stop = true;
    if (conv) st.conv = 1;
    cc[c] = st;
    if (stop) {
        active[c] = 0;
        atomicSub(&gact[cgroup[c]], 1);
        atomicSub(remaining, 1);
    }
}

// The k-th smallest eigenvalue of T (d[n], e[n-1]) in the bracket (lo, hi] by 64-point
// multisection of the Sturm count, down to adjacent doubles; uniform over the wavefront
__device__ inline double wave_kth(int n, const double *d, const double *e, int k, double lo, double hi, int lane)
{
#pragma clang fp contract(off)
    for (int round = 0; round < 64; ++round) {
        const double mid = 0.5 * (lo + hi);
        if (!(mid > lo && mid < hi)) break; // adjacent doubles (uniform)
        const double w = hi - lo;
        const double xj = lo + w * ((double)(lane + 1) / 65.0);
        const unsigned long long above = __ballot(tri_sturm_below(n, d, e, xj) > k);
        if (above == 0ull) {
            lo = lo + w * (64.0 / 65.0);
        } else {
            const int j = __ffsll((long long)above) - 1;
            const double xh = lo + w * ((double)(j + 1) / 65.0), xl = j > 0 ? lo + w * ((double)j / 65.0) : lo;
            lo = xl;
            hi = xh;
        }
    }
    return 0.5 * (lo + hi);
}

// The nev-lowest test of every active copy after a chunk (ed_batch_lowest_k's run): the filter of
// cw_lowest (semi_orth off) on the copy's T_n, with eigenvalues by multisection instead of QL:
// copies within 1e-12 anorm merged; a simple value on an eigenvalue of T^ (T without its first
// row and column: Sturm counts of T^ at v -+ 1e-12 anorm) holds the test back when it has
// converged (until its copy forms) or lies more than 1e-3 anorm above the converged values kept
// below it, and is dropped as a ghost in transit otherwise; every other simple one must pass
// |b_n s[n-1]| < tol anorm (inverse iteration on lane 0).  A copy whose nev lowest all pass, or
// whose recurrence ends in an invariant subspace, or that has run max_steps, is switched off.
// Only when to stop is decided here: the host filters the final tridiagonals again.  Arrays as in
// k_g_check.
template <bool LDS>
__global__ __launch_bounds__(64) void k_g_check_k(int nc, int ld, const double *__restrict__ adev,
                                                  const double *__restrict__ bdev, const int *__restrict__ dchk,
                                                  int chunk, int *active, const int *__restrict__ cgroup, int *gact,
                                                  double tol, int max_steps, int nev, CopyCheck *cc, double *ws,
                                                  char *wsf, int ldz, int *remaining)
{
#pragma clang fp contract(off)
    extern __shared__ double s_chk[];
    const int c = blockIdx.x, lane = threadIdx.x;
    if (c >= nc || !active[c]) return; // uniform over the wavefront
    const int Ne = *dchk + chunk, Nl = Ne - 1;
    double *a = LDS ? s_chk : ws + (size_t)c * 7 * ldz, *b = a + ldz, *w5 = b + ldz;
    char *sw = LDS ? reinterpret_cast<char *>(s_chk + 7 * (size_t)ldz) : wsf + (size_t)c * ldz;
    for (int k = lane; k <= Nl; k += 64) a[k] = adev[(int64_t)k * ld + c];
    for (int k = lane; k <= Ne; k += 64) b[k] = bdev[(int64_t)k * ld + c];
    __syncthreads();
    CopyCheck st = cc[c];
    int M = Nl;
    for (int k = st.processed; k <= Nl; ++k)
        if (!(b[k + 1] > 0.0) || !isfinite(a[k])) {
            M = k;
            break;
        }
    for (int k = st.processed; k <= M; ++k) {
        if (k == 0) st.anorm = fabs(a[0] + b[1]);
        else st.anorm = fmax(st.anorm, b[k] + fabs(a[k] + b[k + 1]));
    }
    st.processed = M + 1;
    const int n = M + 1;
    const double *d = a, *e = b + 1;
    double lo = d[0], hi = d[0];
    for (int i = 0; i < n; ++i) {
        const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < n - 1 ? fabs(e[i]) : 0.0);
        lo = fmin(lo, d[i] - r);
        hi = fmax(hi, d[i] + r);
    }
    const double pad = 1e-12 * (hi - lo + 1.0);
    lo -= pad;
    hi += pad;
    const bool brk = M < Nl || !(b[M + 1] > 0.0);
    const double bres = brk ? 0.0 : b[M + 1];
    const double tm = 1e-12 * st.anorm;
    int found = 0;
    bool ok = true;
    int j = 0;
    double thj = wave_kth(n, d, e, 0, lo, hi, lane), genuine_hi = -INFINITY;
    const double tw = 1e-3 * st.anorm;
    const double th0 = thj;
    while (j < n && found < nev) {
        int r = j + 1;
        double last = thj; // the cluster's largest member
        double thr = r < n ? wave_kth(n, d, e, r, lo, hi, lane) : INFINITY;
        while (r < n && thr - last <= tm) {
            last = thr;
            ++r;
            thr = r < n ? wave_kth(n, d, e, r, lo, hi, lane) : INFINITY;
        }
        bool keep = true, genuine = true;
        if (r - j == 1) {
            int pass = 0;
            if (lane == 0) {
                double *z = w5 + 4 * (size_t)n;
                tri_inverse_iteration(n, d, e, thj, z, w5, sw);
                pass = fabs(bres * z[n - 1]) < tol * st.anorm ? 1 : 0;
            }
            const bool conv1 = __shfl(pass, 0) != 0;
            bool in_that = false;
            if (n > 1) {
                const int cnt = lane < 2 ? tri_sturm_below(n - 1, a + 1, b + 2, lane == 0 ? thj - tm : thj + tm) : 0;
                in_that = __shfl(cnt, 1) > __shfl(cnt, 0);
            }
            bool cv = conv1;
            if (in_that) {
                if (cv) cv = false; // held back until its copy forms
                else keep = thj - genuine_hi > tw;
            }
            if (keep) ok = ok && cv;
            genuine = keep && cv;
        }
        if (keep) ++found;
        if (genuine) genuine_hi = last;
        j = r;
        thj = thr;
    }
    if (lane != 0) return;
    const bool pass = ok && found == nev;
    st.ev = th0;
    st.last = M;
    bool stop = false, conv = false;
    if (brk) {
        stop = true;
        conv = pass;
    } else if (pass) {
        stop = conv = true;
    } else if (Ne >= max_steps) {
        stop = true;
    }
    if (conv) st.conv = 1;
    cc[c] = st;
    if (stop) {
        active[c] = 0;
        atomicSub(&gact[cgroup[c]], 1);
        atomicSub(remaining, 1);
    }
}

// the checks' own step count (the step kernels' counter runs ahead while a check runs beside them)
__global__ void k_g_chk_bump(int *dchk, int chunk) { dchk[0] += chunk; }

// one step: x = f_N, xp = f_{N-1}, y <- f_{N+1}
template <typename V, typename TV>
static int enqueue_step(const Batch &B, const double *x, const double *xp, double *y, double *partial,
                        const int *active, const int *gact, double *adev, double *bdev, int ld, int *dstep,
                        bool first, const double *zdev, const int *lim, double *acc, hipStream_t st)
{
    const int64_t per1 = (B.t1 + 7) / 8, per2 = (B.t2 + 7) / 8;
#define ED_GUP(R_, D_)                                                                                        \
    hipLaunchKernelGGL((k_g_up<V, TV, R_, D_>), dim3((unsigned)(per1 * 8)), dim3(256), B.lds_up, st, B.dgr, B.dtg1, \
                       B.t1, per1, active, gact, bdev, ld, dstep, x, y, B.kpu)
    // DIAG: some group has trailing diagonal UP calls (one load of the row's own amplitude for them)
    if (B.gu == 4) {
        if (B.up_diag) ED_GUP(4, true);
        else ED_GUP(4, false);
    } else {
        if (B.up_diag) ED_GUP(1, true);
        else ED_GUP(1, false);
    }
#undef ED_GUP
    const size_t sh = B.lds_dn;
#define ED_GDN(AX, R_)                                                                                        \
    hipLaunchKernelGGL((k_g_dn<V, TV, AX, R_>), dim3((unsigned)(per2 * 8)), dim3(256), sh, st, B.dgr, B.dtg2, B.t2, \
                       per2, active, gact, bdev, ld, dstep, x, xp, y, partial, B.kp)
    if (B.gr == 8) {
        if (first) ED_GDN(false, 8);
        else ED_GDN(true, 8);
    } else {
        if (first) ED_GDN(false, 4);
        else ED_GDN(true, 4);
    }
#undef ED_GDN
    hipLaunchKernelGGL(k_g_finalize, dim3(B.ncopies), dim3(1024), 0, st, B.dgr, B.dcgroup, active, partial, 0, adev,
                       ld, dstep);
    hipLaunchKernelGGL(k_g_update, dim3((unsigned)B.tc), dim3(256), 0, st, B.dgr, B.dtgc, B.tc, active, gact, adev, bdev,
                       ld, dstep, B.cvec ? 1 : 0, x, y, partial, zdev, lim, acc);
    hipLaunchKernelGGL(k_g_finalize, dim3(B.ncopies), dim3(1024), 0, st, B.dgr, B.dcgroup, active, partial, 1, bdev,
                       ld, dstep);
    hipLaunchKernelGGL(k_g_bump, dim3(1), dim3(1), 0, st, dstep);
    ED_CHECK_LAUNCH();
    return ED_OK;
}

// Ritz vectors by a second run of the same recurrence (same start vectors, kernels and launch
// order, so the same a and b bit for bit): copy c runs lim[c] steps and accumulates
// sum_N z[N * ncopies + c] v_N (z step-major, max_steps + 3 rows); the normalised result goes to
// evecs[instance].
struct BatchRebuild {
    const double *z = nullptr;    // host, step-major, nc per row
    const int *lim = nullptr;     // host, per copy (0: not rebuilt)
    void *const *evecs = nullptr; // per instance, device vectors of the instance's length
};

// The batched solve; ha, hb (optional) get the a and b rows of every copy, step-major (nc per row,
// max_steps + 3 rows; rows past a copy's last step are not set).  rb: the rebuild run instead
// (evals, nsteps, status untouched).
static int batch_run(ed_batch bb, const ed_lanczos_cfg *cfg, uint32_t seed, int use_graph, double *evals,
                     int *nsteps, int *status, void *stream, std::vector<double> *ha, std::vector<double> *hb,
                     const BatchRebuild *rb = nullptr, int nev = 1)
{
    Batch *B = reinterpret_cast<Batch *>(bb);
    if (!B || (!evals && !rb)) return fail(ED_EINVAL, "ed_batch_lowest: bad arguments");
    const int maxN = cfg && cfg->max_steps > 0 ? cfg->max_steps : 1000;
    const double tol = cfg && cfg->tol > 0 ? cfg->tol : 0.01 * std::sqrt(std::numeric_limits<double>::epsilon());
    const int every = 10;
    if (cfg && cfg->check_every > 0 && cfg->check_every != 10)
        return fail(ED_EUNSUPPORTED, "ed_batch_lowest: the batch checks every 10 steps");
    if (cfg && cfg->reorth) return fail(ED_EUNSUPPORTED, "ed_batch_lowest: no reorthogonalisation in the batch");
    // work on a private stream (graph capture needs a non-default stream), ordered after
    // the caller's stream; synchronised before returning
    hipStream_t cst = stream_of(stream), st = nullptr;
    hipEvent_t ev = nullptr;
    ED_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess || hipEventRecord(ev, cst) != hipSuccess ||
        hipStreamWaitEvent(st, ev, 0) != hipSuccess) {
        (void)hipStreamDestroy(st);
        if (ev) (void)hipEventDestroy(ev);
        return fail(ED_EHIP, "ed_batch_lowest: stream setup failed");
    }
    const int nc = B->ncopies, ld = maxN + 3, ldz = maxN + 3;
    // the check kernel keeps its arrays in LDS up to max_steps of about 1070
    constexpr size_t kCheckLdsMax = 60 * 1024;
    const size_t chk_lds = (7 * sizeof(double) + 1) * (size_t)ldz;
    double *buf[3] = {nullptr, nullptr, nullptr}, *partial = nullptr, *adev = nullptr, *bdev = nullptr,
           *tmp = nullptr, *ws = nullptr;
    int *active = nullptr, *dstep = nullptr, *dremaining = nullptr, *dchk = nullptr, *gact = nullptr;
    // the convergence checks run on a side stream, each beside the next chunk's steps
    hipStream_t side = nullptr;
    hipEvent_t ev_chunk = nullptr, ev_check = nullptr, ev_count[2] = {nullptr, nullptr};
    char *wsf = nullptr;
    CopyCheck *cc = nullptr;
    double *dpart = nullptr; // start-vector norm partials + result
    int64_t npart = 0;
    double *zdev = nullptr, *acc = nullptr; // rebuild: Ritz coordinates, accumulated vectors
    int *ldev = nullptr;
    int *pact = nullptr; // pinned: active flags, step counter, copies still running
    // graphs of kGraphChunks chunks (and of one chunk, for the tail) per buffer rotation phase
    constexpr int kGraphChunks = 5;
    hipGraphExec_t gexec[3][3] = {}; // [fifty steps, ten steps, the rest of the solve][phase]
    hipGraph_t graph[3][3] = {};
    int s = ED_OK;
    auto cleanup = [&] {
        (void)hipStreamSynchronize(st);
        for (int q = 0; q < 3; ++q)
            for (int k = 0; k < 3; ++k) {
                if (gexec[q][k]) (void)hipGraphExecDestroy(gexec[q][k]);
                if (graph[q][k]) (void)hipGraphDestroy(graph[q][k]);
            }
        for (int k = 0; k < 3; ++k) dfree(buf[k]);
        dfree(partial);
        dfree(adev);
        dfree(bdev);
        dfree(active);
        dfree(dstep);
        dfree(dremaining);
        dfree(dchk);
        dfree(gact);
        if (side) {
            (void)hipStreamSynchronize(side);
            (void)hipStreamDestroy(side);
        }
        if (ev_chunk) (void)hipEventDestroy(ev_chunk);
        if (ev_check) (void)hipEventDestroy(ev_check);
        for (hipEvent_t e : ev_count)
            if (e) (void)hipEventDestroy(e);
        dfree(tmp);
        dfree(dpart);
        dfree(ws);
        dfree(wsf);
        dfree(cc);
        dfree(zdev);
        dfree(acc);
        dfree(ldev);
        if (pact) (void)hipHostFree(pact);
        (void)hipStreamDestroy(st);
        (void)hipEventDestroy(ev);
    };
    do {
        for (int k = 0; k < 3 && !s; ++k) s = dmalloc((void **)&buf[k], sizeof(double) * (size_t)B->total_real);
        if (s) break;
        int64_t maxe = 1;
        for (const GGroup &g : B->host) maxe = std::max(maxe, g.nelem);
        npart = blas_num_blocks(maxe * (B->cvec ? 2 : 1));
        if ((s = dmalloc((void **)&partial, sizeof(double) * (size_t)std::max<int64_t>(B->pcap, 1))) ||
            (s = dmalloc((void **)&adev, sizeof(double) * (size_t)nc * ld)) ||
            (s = dmalloc((void **)&bdev, sizeof(double) * (size_t)nc * ld)) ||
            (s = dmalloc((void **)&active, sizeof(int) * nc)) || (s = dmalloc((void **)&dstep, sizeof(int) * 2)) ||
            (s = dmalloc((void **)&dremaining, sizeof(int))) || (s = dmalloc((void **)&dchk, sizeof(int))) ||
            (s = dmalloc((void **)&gact, sizeof(int) * (size_t)std::max(B->ng, 1))) ||
            (s = dmalloc((void **)&tmp, sizeof(double) * (size_t)maxe * (B->cvec ? 2 : 1))) ||
            (s = dmalloc((void **)&dpart, sizeof(double) * (size_t)(npart + 1))) ||
            (s = dmalloc((void **)&ws, sizeof(double) * 7 * (size_t)nc * ldz)) ||
            (s = dmalloc((void **)&wsf, (size_t)nc * ldz)) ||
            (s = dmalloc((void **)&cc, sizeof(CopyCheck) * (size_t)nc)))
            break;
        if (rb) {
            if ((s = dmalloc((void **)&zdev, sizeof(double) * (size_t)nc * ld)) ||
                (s = dmalloc((void **)&ldev, sizeof(int) * (size_t)nc)) ||
                (s = dmalloc((void **)&acc, sizeof(double) * (size_t)B->total_real)))
                break;
            if (hipMemcpyAsync(zdev, rb->z, sizeof(double) * (size_t)nc * ld, hipMemcpyHostToDevice, st) != hipSuccess ||
                hipMemcpyAsync(ldev, rb->lim, sizeof(int) * (size_t)nc, hipMemcpyHostToDevice, st) != hipSuccess ||
                hipMemsetAsync(acc, 0, sizeof(double) * (size_t)B->total_real, st) != hipSuccess) {
                s = fail(ED_EHIP, "ed_batch_lowest: rebuild upload failed");
                break;
            }
        }
        if (hipHostMalloc((void **)&pact, sizeof(int) * (size_t)(nc + 3 + B->ng), hipHostMallocDefault) != hipSuccess) {
            s = fail(ED_ENOMEM, "ed_batch_lowest: pinned staging allocation failed");
            break;
        }
        if (hipStreamCreateWithFlags(&side, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&ev_chunk, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&ev_check, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&ev_count[0], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&ev_count[1], hipEventDisableTiming) != hipSuccess) {
            s = fail(ED_EHIP, "ed_batch_lowest: stream setup failed");
            break;
        }
        // start vectors: each copy gets the normalised GenRand vector of its sector length (the
        // padding copies stay zero and inactive); b_0 = |f| as in the single-instance driver,
        // reduced and written into the b_0 row on the device (no host round trip per group)
        const int NC = B->cvec ? 2 : 1;
        for (const GGroup &g : B->host) {
            if (s) break;
            if (hipMemsetAsync(buf[0] + g.voff, 0, sizeof(double) * (size_t)g.nelem * g.G * NC, st) != hipSuccess) {
                s = fail(ED_EHIP, "ed_batch_lowest: clear failed");
                break;
            }
            if ((s = ed_random_vector(g.nelem, seed, B->cvec, tmp, 1, st))) break;
            int64_t nb = 0;
            if ((s = dot_partials(g.nelem * NC, tmp, tmp, dpart, &nb, st)) ||
                (s = finalize_sum(dpart, nb, dpart + npart, 1, st)))
                break;
            hipLaunchKernelGGL(k_g_b0, dim3(1), dim3(64), 0, st, g.cbase, g.ncopy, g.G, (const double *)(dpart + npart),
                               bdev);
            hipLaunchKernelGGL(k_g_scatter, dim3((unsigned)((g.nelem + 255) / 256)), dim3(256), 0, st, g.nelem, g.G,
                               g.ncopy, NC, (const double *)tmp, buf[0] + g.voff);
        }
        if (s) break;
        if (hipGetLastError() != hipSuccess) {
            s = fail(ED_EHIP, "ed_batch_lowest: start vectors failed");
            break;
        }
        int running = 0;
        int *pgact = pact + nc + 3; // running copies per group
        for (int q = 0; q < B->ng; ++q) pgact[q] = 0;
        for (int c = 0; c < nc; ++c) {
            pact[c] = B->inst_of[c] >= 0 && (!rb || rb->lim[c] > 0) ? 1 : 0;
            running += pact[c];
            pgact[B->cgroup[c]] += pact[c];
        }
        pact[nc] = pact[nc + 1] = 0;
        pact[nc + 2] = running;
        if (hipMemcpyAsync(active, pact, sizeof(int) * nc, hipMemcpyHostToDevice, st) ||
            hipMemcpyAsync(dstep, pact + nc, sizeof(int) * 2, hipMemcpyHostToDevice, st) ||
            hipMemcpyAsync(dremaining, pact + nc + 2, sizeof(int), hipMemcpyHostToDevice, st) ||
            hipMemcpyAsync(dchk, pact + nc, sizeof(int), hipMemcpyHostToDevice, st) ||
            hipMemcpyAsync(gact, pgact, sizeof(int) * (size_t)B->ng, hipMemcpyHostToDevice, st) ||
            hipMemsetAsync(cc, 0, sizeof(CopyCheck) * (size_t)nc, st) || hipStreamSynchronize(st)) {
            s = fail(ED_EHIP, "ed_batch_lowest: upload failed");
            break;
        }

        int phase = 0, N = 0;
        auto enqueue = [&](int ph, int n0, int nstep) -> int {
            for (int k = 0; k < nstep; ++k) {
                // buffer rotation: f_N in buf[N % 3], f_{N+1} goes to buf[(N + 1) % 3], f_{N-1}
                // sits in buf[(N + 2) % 3]
                const double *X = buf[(ph + k) % 3], *XP = buf[(ph + k + 2) % 3];
                double *Y = buf[(ph + k + 1) % 3];
                const bool first = (n0 + k) == 0;
                int r = B->cvec ? (B->ctab ? enqueue_step<cplx, cplx>(*B, X, XP, Y, partial, active, gact, adev, bdev,
                                                                     nc, dstep, first, zdev, ldev, acc, st)
                                           : enqueue_step<cplx, double>(*B, X, XP, Y, partial, active, gact, adev,
                                                                       bdev, nc, dstep, first, zdev, ldev, acc, st))
                                : enqueue_step<double, double>(*B, X, XP, Y, partial, active, gact, adev, bdev, nc,
                                                               dstep, first, zdev, ldev, acc, st);
                if (r) return r;
            }
            return ED_OK;
        };
        // nch chunks of ten steps (the last one shorter at max_steps); after each, the
        // device-side convergence test runs on the side stream while the next chunk steps on (a
        // copy found converged is switched off from the chunk after; a kernel that still sees the
        // old flag does wasted but harmless work: a copy's kernels touch only its own vectors,
        // partial slots and a/b rows, and the test reads only the rows of finished steps; each
        // workgroup reads a group's running count once, group_running).  The
        // main stream waits for the last test.  Advances N and the rotation phase.
        auto enqueue_chunks = [&](int nch) -> int {
            for (int q = 0; q < nch && N < maxN; ++q) {
                const int chunk = std::min(every, maxN - N);
                if (int r = enqueue(phase, N, chunk)) return r;
                ED_HIP(hipEventRecord(ev_chunk, st));
                ED_HIP(hipStreamWaitEvent(side, ev_chunk, 0));
                if (rb)
                    hipLaunchKernelGGL(k_g_limit, dim3((unsigned)((nc + 63) / 64)), dim3(64), 0, side, nc,
                                       (const int *)dchk, chunk, active, (const int *)B->dcgroup, gact,
                                       (const int *)ldev, dremaining);
                else if (nev > 1 && chk_lds <= kCheckLdsMax)
                    hipLaunchKernelGGL(k_g_check_k<true>, dim3((unsigned)nc), dim3(64), chk_lds, side, nc, nc,
                                       (const double *)adev, (const double *)bdev, (const int *)dchk, chunk, active,
                                       (const int *)B->dcgroup, gact, tol, maxN, nev, cc, ws, wsf, ldz, dremaining);
                else if (nev > 1)
                    hipLaunchKernelGGL(k_g_check_k<false>, dim3((unsigned)nc), dim3(64), 0, side, nc, nc,
                                       (const double *)adev, (const double *)bdev, (const int *)dchk, chunk, active,
                                       (const int *)B->dcgroup, gact, tol, maxN, nev, cc, ws, wsf, ldz, dremaining);
                else if (chk_lds <= kCheckLdsMax)
                    hipLaunchKernelGGL(k_g_check<true>, dim3((unsigned)nc), dim3(64), chk_lds, side, nc, nc,
                                       (const double *)adev, (const double *)bdev, (const int *)dchk, chunk, active,
                                       (const int *)B->dcgroup, gact, tol, maxN, cc, ws, wsf, ldz, dremaining);
                else
                    hipLaunchKernelGGL(k_g_check<false>, dim3((unsigned)nc), dim3(64), 0, side, nc, nc,
                                       (const double *)adev, (const double *)bdev, (const int *)dchk, chunk, active,
                                       (const int *)B->dcgroup, gact, tol, maxN, cc, ws, wsf, ldz, dremaining);
                hipLaunchKernelGGL(k_g_chk_bump, dim3(1), dim3(1), 0, side, dchk, chunk);
                ED_CHECK_LAUNCH();
                N += chunk;
                phase = (phase + chunk) % 3;
            }
            ED_HIP(hipEventRecord(ev_check, side));
            ED_HIP(hipStreamWaitEvent(st, ev_check, 0));
            return ED_OK;
        };
        // the first chunk holds the step-0 kernel: eagerly.  Then graphs of kGraphChunks chunks
        // (one chunk near max_steps), captured once per phase and replayed, one submission ahead
        // of the host: after each, the count of copies still running is copied into a pinned slot,
        // and the host reads it once the next submission is queued.
        auto submit = [&]() -> int {
            const int left = (maxN - N) / every; // whole chunks left
            // use_graph 2: every whole chunk left in one graph, launched once
            const int q = use_graph == 2 ? 2 : left >= kGraphChunks ? 0 : 1,
                      nch = q == 2 ? left : q == 0 ? kGraphChunks : 1;
            if (!use_graph || left == 0) return enqueue_chunks(left == 0 ? 1 : nch);
            if (!gexec[q][phase]) {
                const int N0 = N, ph0 = phase;
                if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess)
                    return fail(ED_EHIP, "ed_batch_lowest: capture begin failed");
                const int r = enqueue_chunks(nch);
                const hipError_t e = hipStreamEndCapture(st, &graph[q][ph0]);
                N = N0;
                phase = ph0;
                if (r || e != hipSuccess ||
                    hipGraphInstantiate(&gexec[q][ph0], graph[q][ph0], nullptr, nullptr, 0) != hipSuccess)
                    return r ? r : fail(ED_EHIP, "ed_batch_lowest: graph capture failed");
            }
            if (hipGraphLaunch(gexec[q][phase], st) != hipSuccess)
                return fail(ED_EHIP, "ed_batch_lowest: graph launch failed");
            N += nch * every;
            phase = (phase + nch * every) % 3;
            return ED_OK;
        };
        // pinned slots pact[nc], pact[nc + 1] (free after the uploads) take the running counts
        auto post_count = [&](int slot) -> int {
            ED_HIP(hipMemcpyAsync(pact + nc + slot, dremaining, sizeof(int), hipMemcpyDeviceToHost, st));
            ED_HIP(hipEventRecord(ev_count[slot], st));
            return ED_OK;
        };
        if ((s = enqueue_chunks(1)) || (s = post_count(0))) break;
        int pending[2] = {0, 0}, np = 1, next = 1;
        bool stop = false;
        while (np > 0 && !s) {
            if (!stop && N < maxN) {
                if ((s = submit()) || (s = post_count(next))) break;
                pending[np++] = next;
                next ^= 1;
            }
            const int slot = pending[0];
            pending[0] = pending[1];
            --np;
            if (hipEventSynchronize(ev_count[slot]) != hipSuccess) {
                s = fail(ED_EHIP, "ed_batch_lowest: readback failed");
                break;
            }
            if (pact[nc + slot] <= 0) stop = true;
        }
        if (s) break;
        if (rb) {
            // copy -> instance vector, normalised
            for (int i = 0; i < B->n && !s; ++i) {
                const int c = B->copy_of[i];
                if (rb->lim[c] <= 0) continue;
                const GGroup &g = B->host[(size_t)B->cgroup[c]];
                hipLaunchKernelGGL(k_g_gather, dim3((unsigned)((g.nelem + 255) / 256)), dim3(256), 0, st, g.nelem, g.G,
                                   c - g.cbase, NC, (const double *)(acc + g.voff), (double *)rb->evecs[i]);
                if (hipGetLastError() != hipSuccess) s = fail(ED_EHIP, "ed_batch_lowest: gather failed");
                else s = ed_normalize(g.nelem, B->cvec ? 1 : 0, rb->evecs[i], nullptr, st);
            }
            if (!s && hipStreamSynchronize(st) != hipSuccess) s = fail(ED_EHIP, "ed_batch_lowest: rebuild failed");
            break;
        }
        std::vector<CopyCheck> hc((size_t)nc);
        if (hipMemcpyAsync(hc.data(), cc, sizeof(CopyCheck) * (size_t)nc, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            s = fail(ED_EHIP, "ed_batch_lowest: result readback failed");
            break;
        }
        for (int i = 0; i < B->n; ++i) {
            const int c = B->copy_of[i];
            evals[i] = hc[c].ev;
            if (nsteps) nsteps[i] = hc[c].last + 1;
            if (status) status[i] = hc[c].conv ? ED_OK : ED_ENOCONV;
        }
        if (ha && hb) {
            ha->assign((size_t)nc * ld, 0.0);
            hb->assign((size_t)nc * ld, 0.0);
            if (hipMemcpyAsync(ha->data(), adev, sizeof(double) * (size_t)nc * ld, hipMemcpyDeviceToHost, st) !=
                    hipSuccess ||
                hipMemcpyAsync(hb->data(), bdev, sizeof(double) * (size_t)nc * ld, hipMemcpyDeviceToHost, st) !=
                    hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess) {
                s = fail(ED_EHIP, "ed_batch_lowest: tridiagonal readback failed");
                break;
            }
        }
    } while (0);
    cleanup();
    return s;
}

extern "C" int ed_batch_lowest(ed_batch bb, const ed_lanczos_cfg *cfg, uint32_t seed, int use_graph, double *evals,
                               int *nsteps, int *status, void *stream)
{
    return batch_run(bb, cfg, seed, use_graph, evals, nsteps, status, stream, nullptr, nullptr);
}

// The first T_n (n a multiple of ten, or the whole T_M after an invariant subspace) whose lowest
// Ritz pair passes ed_lanczos_eigen's test |b_n s[n-1]| < tol anorm, |s[0]| > tol anorm: the
// step count n at which a single-instance run stops, its eigenvector z and eigenvalue; 0 if none.
static int ground_state_steps(const double *a, const double *b, int M, bool exact, double tol, std::vector<double> &z,
                              double *lam)
{
    std::vector<double> ws, zz;
    std::vector<char> sw;
    double anorm = 0.0;
    int done = 0, n = 0;
    while (n < M) {
        n = std::min(n + 10, M);
        for (int k = done; k < n; ++k)
            anorm = k == 0 ? std::fabs(a[0] + b[1]) : std::fmax(anorm, b[k] + std::fabs(a[k] + b[k + 1]));
        done = n;
        const double l = tri_bisect_kth(n, a, b + 1, 0);
        ws.assign(5 * (size_t)n, 0.0);
        zz.assign((size_t)n, 0.0);
        sw.assign((size_t)n, 0);
        tri_inverse_iteration(n, a, b + 1, l, zz.data(), ws.data(), sw.data());
        if ((exact && n == M) || (std::fabs(b[n] * zz[(size_t)n - 1]) < tol * anorm && std::fabs(zz[0]) > tol * anorm)) {
            z = zz;
            *lam = l;
            return n;
        }
    }
    return 0;
}

extern "C" int ed_batch_lowest_k(ed_batch bb, const ed_lanczos_cfg *cfg, int nev, uint32_t seed, int use_graph,
                                 double *evals, int *nsteps, int *status, void *const *d_evecs, void *stream)
{
    Batch *B = reinterpret_cast<Batch *>(bb);
    if (!B || !evals || nev < 1 || nev > 64) return fail(ED_EINVAL, "ed_batch_lowest_k: bad arguments (nev 1..64)");
    if (cfg && cfg->reorth) return fail(ED_EUNSUPPORTED, "ed_batch_lowest_k: no reorthogonalisation in the batch");
    if (d_evecs)
        for (int i = 0; i < B->n; ++i)
            if (!d_evecs[i]) return fail(ED_EINVAL, "ed_batch_lowest_k: d_evecs[%d] is NULL", i);
    const int maxN = cfg && cfg->max_steps > 0 ? cfg->max_steps : 1000;
    const double tol = cfg && cfg->tol > 0 ? cfg->tol : 0.01 * std::sqrt(std::numeric_limits<double>::epsilon());
    // each copy runs until its nev lowest pass the device test (k_g_check_k), its recurrence ends
    // in an invariant subspace, or max_steps
    const ed_lanczos_cfg run{maxN, tol, 10, 0};
    std::vector<double> ev((size_t)B->n), ha, hb;
    std::vector<int> ns((size_t)B->n), st((size_t)B->n);
    ED_TRY(batch_run(bb, &run, seed, use_graph, ev.data(), ns.data(), st.data(), stream, &ha, &hb, nullptr, nev));
    const int nc = B->ncopies;
    std::vector<double> a, e;
    std::vector<int> lim((size_t)nc, 0);             // rebuild steps per copy
    std::vector<std::vector<double>> zrows((size_t)B->n); // Ritz coordinates per instance
    for (int i = 0; i < B->n; ++i) {
        const int c = B->copy_of[i];
        const int M = ns[(size_t)i]; // steps kept: a[0..M-1], b[0..M]
        a.assign((size_t)M, 0.0);
        e.assign((size_t)std::max(M, 1), 0.0);
        std::vector<double> b((size_t)M + 1);
        for (int k = 0; k < M; ++k) a[(size_t)k] = ha[(size_t)k * nc + c];
        for (int k = 0; k <= M; ++k) b[(size_t)k] = hb[(size_t)k * nc + c];
        double anorm = 0.0;
        for (int k = 0; k < M; ++k)
            anorm = k == 0 ? std::fabs(a[0] + b[1]) : std::fmax(anorm, b[k] + std::fabs(a[k] + b[k + 1]));
        for (int k = 0; k + 1 < M; ++k) e[(size_t)k] = b[(size_t)k + 1];
        const bool exact = !(b[(size_t)M] > 0.0); // the recurrence ended in an invariant subspace
        bool conv = false;
        double *out = evals + (size_t)i * nev;
        const int f = ed::cw_lowest(M, a.data(), e.data(), exact ? 0.0 : b[(size_t)M], nev, tol, anorm, false, out,
                                    &conv);
        for (int k = f; k < nev; ++k) out[k] = std::numeric_limits<double>::quiet_NaN();
        if (nsteps) nsteps[i] = M;
        if (status) status[i] = conv ? ED_OK : ED_ENOCONV;
        if (!d_evecs) continue;
        // the ground state of an instance whose nev pass: from T_n at the step where the lowest
        // pair first passes, whose value must be the filter's lowest
        double lam = 0.0;
        std::vector<double> z;
        const int n0 = conv ? ground_state_steps(a.data(), b.data(), M, exact, tol, z, &lam) : 0;
        if (n0 > 0 && !(std::fabs(lam - out[0]) <= 1e-8 * std::fmax(1.0, std::fabs(out[0])))) {
            if (status) status[i] = ED_ENOCONV;
            continue;
        }
        if (n0 == 0) {
            if (status) status[i] = ED_ENOCONV;
            continue;
        }
        lim[(size_t)c] = n0;
        zrows[(size_t)i] = std::move(z);
    }
    if (!d_evecs) return ED_OK;
    int maxlim = 0;
    for (int l : lim) maxlim = std::max(maxlim, l);
    if (maxlim == 0) return ED_OK;
    const int ld = maxlim + 3;
    std::vector<double> zh((size_t)nc * ld, 0.0);
    for (int i = 0; i < B->n; ++i) {
        const int c = B->copy_of[i];
        for (int k = 0; k < lim[(size_t)c]; ++k) zh[(size_t)k * nc + c] = zrows[(size_t)i][(size_t)k];
    }
    BatchRebuild rb;
    rb.z = zh.data();
    rb.lim = lim.data();
    rb.evecs = d_evecs;
    const ed_lanczos_cfg rcfg{maxlim, 1e-300, 10, 0};
    return batch_run(bb, &rcfg, seed, use_graph, nullptr, nullptr, nullptr, stream, nullptr, nullptr, &rb);
}

extern "C" int ed_batch_set_shape(ed_batch bb, int dn_rows, int dn_panels, int up_rows, int up_panels)
{
    Batch *B = reinterpret_cast<Batch *>(bb);
    if (!B || (dn_rows != 4 && dn_rows != 8) || (up_rows != 1 && up_rows != 4) || dn_panels < 1 || dn_panels > 16 ||
        up_panels < 1 || up_panels > 64)
        return fail(ED_EINVAL, "ed_batch_set_shape: DN rows 4/8, DN panels 1-16, UP rows 1/4, UP panels 1-64");
    B->gr = dn_rows;
    B->kp = dn_panels;
    B->gu = up_rows;
    B->kpu = up_panels;
    return batch_tiles(B);
}

// Batch autotuner: times a fixed-length batched solve (`steps` recurrence steps of every instance,
// no convergence stop) for each panel shape and keeps the fastest.  Stage 1: DN rows (4, 8) x DN
// panels per workgroup (1, 2, 4, 8) with the current UP shape; stage 2: UP rows (1, 4) x UP panels
// (1, 2, 4, 8, 16) with the best DN shape.  The per-element sums and the 4-row units of the dot
// partials do not depend on the shape, so the energies are the same bit for bit.
extern "C" int ed_batch_autotune(ed_batch bb, int steps, ed_batch_tuning *best, void *stream)
{
    Batch *B = reinterpret_cast<Batch *>(bb);
    if (!B || steps < 10) return fail(ED_EINVAL, "ed_batch_autotune: bad arguments (steps >= 10)");
    std::vector<double> ev(B->n);
    std::vector<int> ns(B->n), st(B->n);
    ed_lanczos_cfg cfg{steps, 1e-300, 10, 0};
    auto time_one = [&](double &sec) -> int {
        const auto t0 = std::chrono::steady_clock::now();
        const int s = ed_batch_lowest(bb, &cfg, 235, 1, ev.data(), ns.data(), st.data(), stream);
        sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return s;
    };
    double cand[18];
    for (double &c : cand) c = -1.0;
    double bsec = -1.0;
    int bgr = B->gr, bkp = B->kp, bgu = B->gu, bkpu = B->kpu, ci = 0, s = ED_OK;
    for (int gr : {4, 8})
        for (int kp : {1, 2, 4, 8}) {
            if (s) break;
            if ((s = ed_batch_set_shape(bb, gr, kp, bgu, bkpu))) break;
            double sec = 0.0;
            if ((s = time_one(sec))) break;
            cand[ci++] = sec;
            if (bsec < 0.0 || sec < bsec) {
                bsec = sec;
                bgr = gr;
                bkp = kp;
            }
        }
    for (int gu : {1, 4})
        for (int kpu : {1, 2, 4, 8, 16}) {
            if (s) break;
            if ((s = ed_batch_set_shape(bb, bgr, bkp, gu, kpu))) break;
            double sec = 0.0;
            if ((s = time_one(sec))) break;
            cand[ci++] = sec;
            if (sec < bsec) {
                bsec = sec;
                bgu = gu;
                bkpu = kpu;
            }
        }
    const int s2 = ed_batch_set_shape(bb, bgr, bkp, bgu, bkpu);
    if (s) return s;
    if (s2) return s2;
    if (best) {
        best->dn_rows = bgr;
        best->dn_panels = bkp;
        best->up_rows = bgu;
        best->up_panels = bkpu;
        best->seconds = bsec;
        best->steps = steps;
        for (int i = 0; i < 18; ++i) best->candidate_s[i] = cand[i];
    }
    return ED_OK;
}

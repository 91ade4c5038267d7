// ed_hv_otf.hip -- H.v with hop descriptors generated on the fly (default kernel).
//
// Same operator as Hubbard_mv_ (MBL_test/include/HubMatrix_GPU.cuh:114-161), same
// per-element summation order (UP entries in H_TBq call order, DN entries in call order,
// then U), but the hop lists are not read from ELL tables: for a species state s and an
// operator (i, j) of the H_TBq call list, the neighbour is address[s ^ (1<<i | 1<<j)] and the
// matrix element sign * value with sign = (-1)^popcount(s & between(i,j))
// (fermi_operator.h:100-149).  The reference's ELL slots are filled in call order and each
// nonzero call takes its own slot, so replaying the call list in order visits the nonzero
// slots in slot order.
//
// Why: with tables, the descriptors (12 B per hop, ~19 hops per row/column) are re-read for
// every column slab (UP) and row panel (DN) and cost several times the vector traffic
// (rocprofv3: 145 + 164 GB fetched per 18-site H.v against 38 GB of vector data).
// Generating them costs a few integer ops and one address-table lookup (the 2^N-entry
// rank table, L2-resident) per hop, amortised over W columns (UP) or R rows (DN).
//
// Pass 1 (UP term) -- tiles of 256/TPR rows x W = TPR*VEC columns (a column slab), slabs
// split over the 8 XCDs, each XCD walking its slabs one at a time.  For each row the 64
// lanes of a wavefront evaluate up to 64 operator calls in parallel, a ballot orders the
// hopping ones, and the (column, value) list lands in LDS.  Then lane (row, col) streams
// x[neighbour row, col] over the list.  Large sectors visit the rows of a slab in reverse
// Cuthill-McKee order of the hopping graph, so a row's neighbour rows were mostly read
// shortly before (L2 hits) instead of being fetched again.
// Pass 2 (DN term + diagonal + Lanczos epilogue) -- tiles of R rows x 256 columns,
// XCD-aware row-panel order; each lane walks the call list for its column once and
// applies every hop to its R rows, starting from the pass-1 partial.
#include <algorithm>
#include <cmath>
#include <vector>

#include "ed_common.hpp"
#include "ed_hubbard.hpp"
#include "ed_hv_otf_body.hpp"

namespace ed {

template <typename V, typename TV, bool SCALE, int TPR, int VEC>
__global__ __launch_bounds__(256) void k_otf_up(OtfArgs p, int64_t ntiles, int64_t per_xcd, int64_t nrp,
                                                const V *__restrict__ x, V *out, const double *__restrict__ bdev,
                                                int step)
{
    constexpr int ORB = 256 / TPR;
    extern __shared__ double smem_o[];
    const int nops = p.up.nops;
    TV *s_tab = reinterpret_cast<TV *>(smem_o);                       // [4 nops] (entry table only)
    TV *s_val = s_tab + 4 * nops;                                     // [ORB][nops]
    int32_t *s_col = reinterpret_cast<int32_t *>(s_val + ORB * nops); // [ORB][nops]
    __shared__ int s_cnt[ORB];
    __shared__ int32_t s_row[ORB];
    const int64_t b = blockIdx.x;
    const int64_t t = per_xcd ? (b & 7) * per_xcd + (b >> 3) : b; // per_xcd = 0: plain order
    if (t >= ntiles) return; // uniform over the workgroup
    const double sc = SCALE ? 1.0 / bdev[step] : 1.0;
    otf_up_body<V, TV, SCALE, TPR, VEC>(p, t, nrp, x, out, sc, s_val, s_col, s_cnt, s_row, s_tab);
}

template <typename V, typename TV, int R, int FLAGS, bool MASKED>
__global__ __launch_bounds__(256) void k_otf_dn(OtfArgs p, int64_t ntiles, int64_t per_xcd, int64_t ncb, bool has_up,
                                                const V *__restrict__ x, const V *ypart, V *y, V *vout, const V *w,
                                                const double *__restrict__ bdev, int step,
                                                double *__restrict__ partial)
{
    __shared__ double s_red[4];
    const int64_t b = blockIdx.x;
    const int64_t t = (b & 7) * per_xcd + (b >> 3);
    if (t >= ntiles) return; // uniform over the workgroup
    const double sc = (FLAGS & HV_SCALE) ? 1.0 / bdev[step] : 1.0;
    const double c = (FLAGS & HV_AXPY) ? -bdev[step] : 0.0;
    const double r = otf_dn_body<V, TV, R, FLAGS, MASKED>(p, t, ncb, has_up, x, ypart, y, vout, w, sc, c, s_red);
    if ((FLAGS & HV_DOT) && threadIdx.x == 0) partial[t] = r;
}

// Pass 2, descriptor-cached form: a workgroup owns 256 columns and a group of K row panels
// of R rows.  Each lane first writes its column's element of every call into LDS as
// (column << 8 | code), code = 4 * call + {+v_f, -v_f, +v_b, -v_b} (all ones: no element),
// with the values in a 4 * nops table; then it walks the K panels, gathering only active hops.  The descriptors
// are built once per K*R rows while a panel stays R rows tall, so the rows a workgroup has
// in flight (and the XCD's L2 working set) stay small.  Needs dim_dn < 2^24, nops <= 64.
template <typename V, typename TV, int R, int K, int FLAGS>
__global__ __launch_bounds__(256) void k_otf_dn2(OtfArgs p, int64_t ntiles, int64_t per_xcd, int64_t ncb, bool has_up,
                                                 const V *__restrict__ x, const V *ypart, V *y, V *vout, const V *w,
                                                 const double *__restrict__ bdev, int step,
                                                 double *__restrict__ partial)
{
    extern __shared__ double smem_d[];
    __shared__ double s_red[4];
    const int nops = p.dn.nops;
    TV *s_tab = reinterpret_cast<TV *>(smem_d);                      // [4 * nops]
    uint32_t *s_desc = reinterpret_cast<uint32_t *>(s_tab + 4 * nops); // [nops][256]
    const int64_t b = blockIdx.x;
    const int64_t t = (b & 7) * per_xcd + (b >> 3);
    if (t >= ntiles) return; // uniform over the workgroup
    const double sc = (FLAGS & HV_SCALE) ? 1.0 / bdev[step] : 1.0;
    const double c = (FLAGS & HV_AXPY) ? -bdev[step] : 0.0;
    const int64_t grp = t / ncb, cb = t - grp * ncb;
    const int64_t dim_dn = p.dim_dn;
    const int64_t d = cb * 256 + threadIdx.x;
    for (int q = threadIdx.x; q < 4 * nops; q += 256) {
        const OtfOp op = p.dn.ops[q >> 2];
        const int k = q & 3;
        const TV v = (k < 2) ? tv_make<TV>(op.vf_re, op.vf_im) : tv_make<TV>(op.vb_re, op.vb_im);
        s_tab[q] = (k & 1) ? tv_neg<TV>(v) : v;
    }
    uint32_t sd = 0;
    if (d < dim_dn) sd = (uint32_t)p.state_dn[d];
    for (int o = 0; o < nops; ++o) {
        uint32_t desc = 0xffffffffu; // no element of call o in this column
        if (d < dim_dn) {
            const OtfOp op = p.dn.ops[o];
            const uint32_t oi = (sd >> op.i) & 1u, oj = (sd >> op.j) & 1u;
            if (op.kind == ED_OP_DENSITY) {
                if (oi) desc = ((uint32_t)d << 8) | (uint32_t)(4 * o);
            } else if (op.kind == ED_OP_DENSITY2) {
                if (oi & oj) desc = ((uint32_t)d << 8) | (uint32_t)(4 * o);
            } else if (oi != oj) {
                const uint32_t code = (oi ? 0u : 2u) | (uint32_t)(__popc(sd & op.between) & 1);
                desc = ((uint32_t)p.dn.addr[sd ^ op.m] << 8) | (uint32_t)(4 * o) | code;
            }
        }
        s_desc[o * 256 + threadIdx.x] = desc;
    }
    __syncthreads();
    double dot = 0.0;
    for (int pn = 0; pn < K; ++pn) {
        const int64_t u0 = (grp * K + pn) * R;
        if (u0 >= p.dim_up) break; // uniform
        if (d < dim_dn) {
            V acc[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int64_t u = u0 + r;
                acc[r] = (has_up && u < p.dim_up) ? ypart[u * dim_dn + d] : VO<V>::zero();
            }
            // calls in order, uniform over the wave (lanes gather for the same call, whose
            // targets are nearly contiguous); lanes without an element of the call do not load
            constexpr int KB = 4;
            for (int o0 = 0; o0 < nops; o0 += KB) {
                uint32_t dk[KB];
                V xv[KB][R];
#pragma unroll
                for (int k = 0; k < KB; ++k) {
                    dk[k] = (o0 + k < nops) ? s_desc[(o0 + k) * 256 + threadIdx.x] : 0xffffffffu;
                    if (dk[k] != 0xffffffffu) {
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            const int64_t u = (u0 + r < p.dim_up) ? u0 + r : u0;
                            xv[k][r] = x[u * dim_dn + (dk[k] >> 8)];
                        }
                    }
                }
#pragma unroll
                for (int k = 0; k < KB; ++k) {
                    if (dk[k] != 0xffffffffu) {
                        const TV a = s_tab[dk[k] & 0xffu];
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            V xk = xv[k][r];
                            if (FLAGS & HV_SCALE) xk = VO<V>::scale(sc, xk);
                            acc[r] = omac(a, xk, acc[r]);
                        }
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int64_t u = u0 + r;
                if (u >= p.dim_up) continue;
                const int64_t i = u * dim_dn + d;
                V xd = x[i];
                if (FLAGS & HV_SCALE) xd = VO<V>::scale(sc, xd);
                if (p.has_diag) acc[r] = VO<V>::fmr(diag_of(p.state_up[u], (int32_t)sd, p), xd, acc[r]);
                V o = acc[r];
                if (FLAGS & HV_AXPY) o = VO<V>::fmr(c, w[i], o);
                y[i] = o;
                if (FLAGS & HV_WRITEV) vout[i] = xd;
                if (FLAGS & HV_DOT) dot += VO<V>::dotre(xd, o);
            }
        }
    }
    if (FLAGS & HV_DOT) {
        const double r = block_sum<256>(dot, s_red);
        if (threadIdx.x == 0) partial[t] = r;
    }
}

// Pass-1 entry table, one thread per up row u: the row's (column << 8 | value code) words in
// call order (the order op_entry visits them), then 0xffffffff padding up to nops words.
__global__ __launch_bounds__(256) void k_up_entries(OtfArgs p, uint32_t *__restrict__ tab, int32_t *__restrict__ cnt)
{
    const int64_t u = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (u >= p.dim_up) return;
    const int nops = p.up.nops;
    const uint32_t s = (uint32_t)p.state_up[u];
    uint32_t *row = tab + u * nops;
    int n = 0;
    for (int o = 0; o < nops; ++o) {
        const OtfOp op = p.up.ops[o];
        const uint32_t oi = (s >> op.i) & 1u, oj = (s >> op.j) & 1u;
        uint32_t e = 0xffffffffu;
        if (op.kind == ED_OP_DENSITY) {
            if (oi) e = ((uint32_t)u << 8) | (uint32_t)(4 * o);
        } else if (op.kind == ED_OP_DENSITY2) {
            if (oi & oj) e = ((uint32_t)u << 8) | (uint32_t)(4 * o);
        } else if (oi != oj) {
            const uint32_t code = (oi ? 0u : 2u) | (uint32_t)(__popc(s & op.between) & 1);
            e = ((uint32_t)p.up.addr[s ^ op.m] << 8) | (uint32_t)(4 * o) | code;
        }
        if (e != 0xffffffffu) row[n++] = e;
    }
    for (int j = n; j < nops; ++j) row[j] = 0xffffffffu;
    cnt[u] = n;
}

int otf_build_up_table(Hubbard &h, hipStream_t st)
{
    const OtfArgs &p = h.otf;
    if (p.up.nops == 0 || p.up.nops > 64 || p.dim_up >= ((int64_t)1 << 24)) return ED_OK; // generated on the fly
    ED_TRY(dmalloc((void **)&h.d_up_tab, sizeof(uint32_t) * (size_t)p.dim_up * p.up.nops));
    ED_TRY(dmalloc((void **)&h.d_up_cnt, sizeof(int32_t) * (size_t)p.dim_up));
    hipLaunchKernelGGL(k_up_entries, dim3((unsigned)((p.dim_up + 255) / 256)), dim3(256), 0, st, p, h.d_up_tab,
                       h.d_up_cnt);
    ED_CHECK_LAUNCH();
    ED_HIP(hipStreamSynchronize(st));
    return ED_OK;
}

struct Plan3 {
    int R, W, TPR, VEC, K;
    bool dn2;
    int64_t nslab, nrp1, tiles1, per1, ncb, nrp2, tiles2, per2;
};

static Plan3 plan3(const Hubbard &h)
{
    Plan3 P;
    const int64_t du = h.args.dim_up, dd = h.args.dim_dn;
    // slab width W = TPR threads per row x VEC columns per thread; VEC = 2 (16 B loads per lane)
    // needs an even dim_dn so that every row chunk stays 16 B aligned
    int w = 8;
    for (int c = 16; c <= 512; c *= 2)
        if (h.hv_w >= c) w = c;
    P.VEC = (w >= 16 && dd % 2 == 0) ? 2 : 1;
    P.TPR = std::min(w / P.VEC, 256);
    P.W = P.TPR * P.VEC;
    P.nslab = (dd + P.W - 1) / P.W;
    P.nrp1 = (du + 256 / P.TPR - 1) / (256 / P.TPR);
    P.tiles1 = P.nslab * P.nrp1;
    P.per1 = (P.tiles1 + 7) / 8;
    P.ncb = (dd + 255) / 256;
    P.R = 1;
    for (int R : {8, 4, 2}) {
        if (R <= h.hv_rmax && ((du + R - 1) / R) * P.ncb >= 4096) {
            P.R = R;
            break;
        }
    }
    // descriptor-cached pass 2: panels of 4 rows, 16 panels per workgroup
    const size_t tv = h.cplx_tables ? 16 : 8;
    P.dn2 = h.hv_dn2 && dd < (1 << 24) && h.otf.dn.nops <= 64 &&
            4 * h.otf.dn.nops * tv + (size_t)h.otf.dn.nops * 256 * 4 <= 65536;
    P.K = 1;
    if (P.dn2) {
        P.R = (h.hv_rmax >= 8) ? 8 : 4;
        P.K = (h.hv_dnk == 1 || h.hv_dnk == 2 || h.hv_dnk == 4 || h.hv_dnk == 16) ? h.hv_dnk : 8;
    }
    P.nrp2 = P.dn2 ? (du + P.R * P.K - 1) / (P.R * P.K) : (du + P.R - 1) / P.R;
    P.tiles2 = P.nrp2 * P.ncb;
    P.per2 = (P.tiles2 + 7) / 8;
    return P;
}

int64_t otf_num_blocks(const Hubbard &h) { return plan3(h).tiles2; }

template <typename V, typename TV, int FLAGS>
static int launch_otf(const Hubbard &h, const V *x, V *y, V *vout, const V *w, const double *bdev, int step,
                      double *partial, hipStream_t st, int64_t *nblocks_out)
{
    const Plan3 P = plan3(h);
    if (P.per1 * 8 > 0x7fffffffLL || P.per2 * 8 > 0x7fffffffLL) return fail(ED_EUNSUPPORTED, "H.v grid too large");
    const bool has_up = h.otf.up.nhop > 0;
    V *upbuf = (FLAGS & HV_WRITEV) ? vout : y;
    constexpr bool SC = (FLAGS & HV_SCALE) != 0;
    if (has_up) {
        const size_t sh = (size_t)(256 / P.TPR + 4) * h.otf.up.nops * sizeof(TV) +
                          (size_t)(256 / P.TPR) * h.otf.up.nops * sizeof(int32_t);
        // XCD-partitioned slabs (each XCD walks its own range of slabs) or, with hv_upxcd = 0,
        // all XCDs on the same slabs at once (one slab's rows in MALL instead of eight)
        const int64_t pper = h.hv_upxcd ? P.per1 : 0;
        const dim3 g1((unsigned)(h.hv_upxcd ? P.per1 * 8 : P.tiles1));
        OtfArgs oa = h.otf;
        oa.up_order = h.hv_uporder ? h.d_up_order : nullptr;
        oa.up_tab = h.hv_uptab ? h.d_up_tab : nullptr;
        oa.up_cnt = h.d_up_cnt;
#define ED_UP(TPR_, VEC_)                                                                                \
    hipLaunchKernelGGL((k_otf_up<V, TV, SC, TPR_, VEC_>), g1, dim3(256), sh, st, oa, P.tiles1, pper, P.nrp1, x, \
                       upbuf, bdev, step)
        if (P.VEC == 2) {
            switch (P.TPR) {
            case 256: ED_UP(256, 2); break;
            case 128: ED_UP(128, 2); break;
            case 64: ED_UP(64, 2); break;
            case 32: ED_UP(32, 2); break;
            case 16: ED_UP(16, 2); break;
            default: ED_UP(8, 2); break;
            }
        } else {
            switch (P.TPR) {
            case 256: ED_UP(256, 1); break;
            case 128: ED_UP(128, 1); break;
            case 64: ED_UP(64, 1); break;
            case 32: ED_UP(32, 1); break;
            case 16: ED_UP(16, 1); break;
            default: ED_UP(8, 1); break;
            }
        }
#undef ED_UP
        ED_CHECK_LAUNCH();
    }
    dim3 g2((unsigned)(P.per2 * 8)), b2(256);
    if (P.dn2) {
        const size_t sh2 = (size_t)4 * h.otf.dn.nops * sizeof(TV) + (size_t)h.otf.dn.nops * 256 * sizeof(uint32_t);
#define ED_DN2(R_, K_)                                                                                          \
    hipLaunchKernelGGL((k_otf_dn2<V, TV, R_, K_, FLAGS>), g2, b2, sh2, st, h.otf, P.tiles2, P.per2, P.ncb, has_up, x, \
                       upbuf, y, vout, w, bdev, step, partial)
        if (P.R == 8) {
            switch (P.K) {
            case 1: ED_DN2(8, 1); break;
            case 2: ED_DN2(8, 2); break;
            case 4: ED_DN2(8, 4); break;
            case 16: ED_DN2(8, 16); break;
            default: ED_DN2(8, 8); break;
            }
        } else {
            switch (P.K) {
            case 1: ED_DN2(4, 1); break;
            case 2: ED_DN2(4, 2); break;
            case 4: ED_DN2(4, 4); break;
            case 16: ED_DN2(4, 16); break;
            default: ED_DN2(4, 8); break;
            }
        }
#undef ED_DN2
        ED_CHECK_LAUNCH();
        if (nblocks_out) *nblocks_out = P.tiles2;
        return ED_OK;
    }
#define ED_DN(R_, M_)                                                                                          \
    hipLaunchKernelGGL((k_otf_dn<V, TV, R_, FLAGS, M_>), g2, b2, 0, st, h.otf, P.tiles2, P.per2, P.ncb, has_up, x, \
                       upbuf, y, vout, w, bdev, step, partial)
    if (h.hv_dnmask) {
        switch (P.R) {
        case 8: ED_DN(8, true); break;
        case 4: ED_DN(4, true); break;
        case 2: ED_DN(2, true); break;
        default: ED_DN(1, true); break;
        }
    } else {
        switch (P.R) {
        case 8: ED_DN(8, false); break;
        case 4: ED_DN(4, false); break;
        case 2: ED_DN(2, false); break;
        default: ED_DN(1, false); break;
        }
    }
#undef ED_DN
    ED_CHECK_LAUNCH();
    if (nblocks_out) *nblocks_out = P.tiles2;
    return ED_OK;
}

template <int FLAGS>
int otf_dispatch(const Hubbard &h, bool cvec, const void *x, void *y, void *vout, const void *w, const double *bdev,
                 int step, double *partial, hipStream_t st, int64_t *nblocks)
{
    if (!cvec)
        return launch_otf<double, double, FLAGS>(h, (const double *)x, (double *)y, (double *)vout,
                                                 (const double *)w, bdev, step, partial, st, nblocks);
    if (h.cplx_tables)
        return launch_otf<cplx, cplx, FLAGS>(h, (const cplx *)x, (cplx *)y, (cplx *)vout, (const cplx *)w, bdev,
                                             step, partial, st, nblocks);
    return launch_otf<cplx, double, FLAGS>(h, (const cplx *)x, (cplx *)y, (cplx *)vout, (const cplx *)w, bdev, step,
                                           partial, st, nblocks);
}

template int otf_dispatch<0>(const Hubbard &, bool, const void *, void *, void *, const void *, const double *, int,
                             double *, hipStream_t, int64_t *);
template int otf_dispatch<HV_SCALE | HV_WRITEV | HV_DOT>(const Hubbard &, bool, const void *, void *, void *,
                                                         const void *, const double *, int, double *, hipStream_t,
                                                         int64_t *);
template int otf_dispatch<HV_SCALE | HV_WRITEV | HV_AXPY | HV_DOT>(const Hubbard &, bool, const void *, void *, void *,
                                                                   const void *, const double *, int, double *,
                                                                   hipStream_t, int64_t *);
template int otf_dispatch<HV_SCALE | HV_WRITEV>(const Hubbard &, bool, const void *, void *, void *, const void *,
                                                const double *, int, double *, hipStream_t, int64_t *);
template int otf_dispatch<HV_SCALE | HV_WRITEV | HV_AXPY>(const Hubbard &, bool, const void *, void *, void *,
                                                          const void *, const double *, int, double *, hipStream_t,
                                                          int64_t *);

// ---------------------------------------------------------------------------------
// operator-list preparation (values computed on the device with the same sincos as the
// table builder, so both paths use bit-identical matrix elements)
// ---------------------------------------------------------------------------------
__global__ void k_prep_ops(int n, const ed_op *__restrict__ in, OtfOp *__restrict__ out)
{
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n) return;
    const ed_op a = in[o];
    OtfOp r;
    r.kind = a.kind;
    r.i = a.i;
    r.j = a.j;
    r.m = (1u << a.i) | (1u << a.j);
    const int lo = a.i < a.j ? a.i : a.j, hi = a.i < a.j ? a.j : a.i;
    r.between = (hi - lo > 1) ? ((((1u << (hi - lo - 1)) - 1u)) << (lo + 1)) : 0u;
    r.vf_im = r.vb_im = 0.0;
    switch (a.kind) {
    case ED_OP_HOP: r.vf_re = r.vb_re = a.value; break;
    case ED_OP_CURRENT: // row with i empty, j occupied gets +t (c+_i c_j), the other -t
        r.vb_re = a.value;
        r.vf_re = -a.value;
        break;
    case ED_OP_HOP_PHASE: {
        double s, c;
        sincos(a.phase, &s, &c);
        r.vf_re = a.value * c;
        r.vf_im = a.value * s;
        sincos(-a.phase, &s, &c);
        r.vb_re = a.value * c;
        r.vb_im = a.value * s;
        break;
    }
    default: // DENSITY / DENSITY2
        r.vf_re = r.vb_re = a.value;
        break;
    }
    out[o] = r;
}

int otf_prepare(OtfSpecies &sp, const ed_op *ops, int nops, const int32_t *d_address, hipStream_t st)
{
    sp.nops = nops;
    sp.addr = d_address;
    int nhop = 0;
    for (int o = 0; o < nops; ++o) nhop += 1;
    sp.nhop = nhop;
    ED_TRY(dmalloc((void **)&sp.ops_own, sizeof(OtfOp) * std::max(1, nops)));
    sp.ops = sp.ops_own;
    if (nops == 0) return ED_OK;
    ed_op *din = nullptr;
    ED_TRY(dmalloc((void **)&din, sizeof(ed_op) * nops));
    int s = ED_OK;
    if (hipMemcpyAsync(din, ops, sizeof(ed_op) * nops, hipMemcpyHostToDevice, st) != hipSuccess)
        s = fail(ED_EHIP, "otf_prepare: upload failed");
    if (!s) {
        hipLaunchKernelGGL(k_prep_ops, dim3((nops + 63) / 64), dim3(64), 0, st, nops, din, sp.ops_own);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
            s = fail(ED_EHIP, "otf_prepare: kernel failed");
    }
    dfree(din);
    return s;
}

} // namespace ed

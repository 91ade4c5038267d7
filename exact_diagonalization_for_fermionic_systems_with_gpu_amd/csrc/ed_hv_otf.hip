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
// Pass 1 (UP term) -- tiles of 32 rows x 8 columns, XCD-aware slab-major order.  Each
// wavefront owns 8 rows; for each row the 64 lanes evaluate up to 64 operator calls in
// parallel, a ballot orders the hopping ones, and the (column, value) list lands in LDS.
// Then lane (row, col) streams x[neighbour row, col] over the list.
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

template <typename V, typename TV, bool SCALE, int OW>
__global__ __launch_bounds__(256) void k_otf_up(OtfArgs p, int64_t ntiles, int64_t per_xcd, int64_t nrp,
                                                const V *__restrict__ x, V *out, const double *__restrict__ bdev,
                                                int step)
{
    constexpr int ORB = 256 / OW;
    extern __shared__ double smem_o[];
    const int nops = p.up.nops;
    TV *s_val = reinterpret_cast<TV *>(smem_o);                       // [ORB][nops]
    int32_t *s_col = reinterpret_cast<int32_t *>(s_val + ORB * nops); // [ORB][nops]
    __shared__ int s_cnt[ORB];
    const int64_t b = blockIdx.x;
    const int64_t t = (b & 7) * per_xcd + (b >> 3);
    if (t >= ntiles) return; // uniform over the workgroup
    const double sc = SCALE ? 1.0 / bdev[step] : 1.0;
    otf_up_body<V, TV, SCALE, OW>(p, t, nrp, x, out, sc, s_val, s_col, s_cnt);
}

template <typename V, typename TV, int R, int FLAGS>
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
    const double r = otf_dn_body<V, TV, R, FLAGS>(p, t, ncb, has_up, x, ypart, y, vout, w, sc, c, s_red);
    if ((FLAGS & HV_DOT) && threadIdx.x == 0) partial[t] = r;
}

struct Plan3 {
    int R, W;
    int64_t nslab, nrp1, tiles1, per1, ncb, nrp2, tiles2, per2;
};

static Plan3 plan3(const Hubbard &h)
{
    Plan3 P;
    const int64_t du = h.args.dim_up, dd = h.args.dim_dn;
    P.W = (h.hv_w == 16 || h.hv_w == 32 || h.hv_w == 64) ? h.hv_w : 8;
    P.nslab = (dd + P.W - 1) / P.W;
    P.nrp1 = (du + 256 / P.W - 1) / (256 / P.W);
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
    P.nrp2 = (du + P.R - 1) / P.R;
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
        const size_t sh = (size_t)(256 / P.W) * h.otf.up.nops * (sizeof(TV) + sizeof(int32_t));
        const dim3 g1((unsigned)(P.per1 * 8));
        switch (P.W) {
        case 64:
            hipLaunchKernelGGL((k_otf_up<V, TV, SC, 64>), g1, dim3(256), sh, st, h.otf, P.tiles1, P.per1, P.nrp1, x,
                               upbuf, bdev, step);
            break;
        case 32:
            hipLaunchKernelGGL((k_otf_up<V, TV, SC, 32>), g1, dim3(256), sh, st, h.otf, P.tiles1, P.per1, P.nrp1, x,
                               upbuf, bdev, step);
            break;
        case 16:
            hipLaunchKernelGGL((k_otf_up<V, TV, SC, 16>), g1, dim3(256), sh, st, h.otf, P.tiles1, P.per1, P.nrp1, x,
                               upbuf, bdev, step);
            break;
        default:
            hipLaunchKernelGGL((k_otf_up<V, TV, SC, 8>), g1, dim3(256), sh, st, h.otf, P.tiles1, P.per1, P.nrp1, x,
                               upbuf, bdev, step);
            break;
        }
        ED_CHECK_LAUNCH();
    }
    dim3 g2((unsigned)(P.per2 * 8)), b2(256);
    switch (P.R) {
    case 8:
        hipLaunchKernelGGL((k_otf_dn<V, TV, 8, FLAGS>), g2, b2, 0, st, h.otf, P.tiles2, P.per2, P.ncb, has_up, x,
                           upbuf, y, vout, w, bdev, step, partial);
        break;
    case 4:
        hipLaunchKernelGGL((k_otf_dn<V, TV, 4, FLAGS>), g2, b2, 0, st, h.otf, P.tiles2, P.per2, P.ncb, has_up, x,
                           upbuf, y, vout, w, bdev, step, partial);
        break;
    case 2:
        hipLaunchKernelGGL((k_otf_dn<V, TV, 2, FLAGS>), g2, b2, 0, st, h.otf, P.tiles2, P.per2, P.ncb, has_up, x,
                           upbuf, y, vout, w, bdev, step, partial);
        break;
    default:
        hipLaunchKernelGGL((k_otf_dn<V, TV, 1, FLAGS>), g2, b2, 0, st, h.otf, P.tiles2, P.per2, P.ncb, has_up, x,
                           upbuf, y, vout, w, bdev, step, partial);
        break;
    }
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

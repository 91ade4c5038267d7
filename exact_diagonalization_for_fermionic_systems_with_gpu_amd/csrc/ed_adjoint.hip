// ed_adjoint.hip -- bilinear forms <a| dH/dtheta |b> for every operator call of a species and
// for every site's double occupancy: the parameter part of the adjoint of y = H(theta) x and
// the Hellmann-Feynman gradients dE0/dtheta = <psi| dH/dtheta |psi>.
//
// H(theta) is linear in each call's value, so for call o of a species (value t_o, phase phi_o,
// fermi_operator.h:100-149 with the Peierls extension) the two directional pieces
//   F_o = sum_{u,d} conj(a[u,d]) H^{o,i<-j}_{u,u'} b[u',d]    (rows with site i occupied)
//   G_o = sum_{u,d} conj(a[u,d]) H^{o,j<-i}_{u,u'} b[u',d]    (rows with site j occupied)
// (for the down species the roles of u and d swap) give dH/dt_o = (F_o + G_o)/t_o and
// dH/dphi_o = i (F_o - G_o).  Density calls only have F.  The double occupancy of site i is
// sum over (u,d) with bit i set in u & d of conj(a) b.  The hop lists are generated exactly
// as in the on-the-fly H.v (address-table ranks, popcount signs).  One lane per element,
// per-call sums in registers, deterministic block and grid reductions.
#include <algorithm>
#include <vector>

#include "ed_common.hpp"
#include "ed_hubbard.hpp"
#include "ed_hv_otf_body.hpp"

namespace ed {

constexpr int kMaxBondOps = 16; // calls per species handled by one kernel launch

// out[(blk * 2*nops + 2*o + dir) * 2 + {re,im}]: per-block partials
template <bool CPLX, bool DNSIDE>
__global__ __launch_bounds__(256) void k_bond_bilinear(OtfArgs p, int o0, int no, const double *__restrict__ a,
                                                       const double *__restrict__ b, double *__restrict__ partial)
{
    __shared__ double red[4];
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t D = p.dim_up * p.dim_dn;
    double fr[kMaxBondOps], fi[kMaxBondOps], gr[kMaxBondOps], gi[kMaxBondOps];
#pragma unroll
    for (int k = 0; k < kMaxBondOps; ++k) fr[k] = fi[k] = gr[k] = gi[k] = 0.0;
    if (g < D) {
        const int64_t u = g / p.dim_dn, d = g - u * p.dim_dn;
        const OtfSpecies &S = DNSIDE ? p.dn : p.up;
        const uint32_t s = DNSIDE ? (uint32_t)p.state_dn[d] : (uint32_t)p.state_up[u];
        const int32_t self = DNSIDE ? (int32_t)d : (int32_t)u;
        double ar, ai = 0.0;
        if (CPLX) {
            ar = a[2 * g];
            ai = a[2 * g + 1];
        } else {
            ar = a[g];
        }
#pragma unroll
        for (int k = 0; k < kMaxBondOps; ++k) {
            if (k >= no) break;
            const OtfOp op = S.ops[o0 + k];
            int32_t col;
            cplx v;
            if (!op_entry<cplx>(op, s, self, S.addr, col, v)) continue;
            const int64_t src = DNSIDE ? u * p.dim_dn + col : (int64_t)col * p.dim_dn + d;
            double br, bi = 0.0;
            if (CPLX) {
                br = b[2 * src];
                bi = b[2 * src + 1];
            } else {
                br = b[src];
            }
            // conj(a) * v * b
            const double vbr = v.re * br - v.im * bi, vbi = v.re * bi + v.im * br;
            const double tr = ar * vbr + ai * vbi, ti = ar * vbi - ai * vbr;
            const bool fwd = op.kind == ED_OP_DENSITY || op.kind == ED_OP_DENSITY2 || ((s >> op.i) & 1u);
            if (fwd) {
                fr[k] += tr;
                fi[k] += ti;
            } else {
                gr[k] += tr;
                gi[k] += ti;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kMaxBondOps; ++k) {
        if (k >= no) break; // uniform
        const double vals[4] = {fr[k], fi[k], gr[k], gi[k]};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double r = block_sum<256>(vals[q], red);
            if (threadIdx.x == 0) partial[((int64_t)blockIdx.x * no + k) * 4 + q] = r;
        }
    }
}

template <bool CPLX>
__global__ __launch_bounds__(256) void k_site_docc(OtfArgs p, int N, const double *__restrict__ a,
                                                   const double *__restrict__ b, double *__restrict__ partial)
{
    __shared__ double red[4];
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t D = p.dim_up * p.dim_dn;
    double acc[32];
#pragma unroll
    for (int i = 0; i < 32; ++i) acc[i] = 0.0;
    if (g < D) {
        const int64_t u = g / p.dim_dn, d = g - u * p.dim_dn;
        const uint32_t dbl = (uint32_t)(p.state_up[u] & p.state_dn[d]);
        double re;
        if (CPLX) re = a[2 * g] * b[2 * g] + a[2 * g + 1] * b[2 * g + 1];
        else re = a[g] * b[g];
#pragma unroll
        for (int i = 0; i < 32; ++i)
            if ((dbl >> i) & 1u) acc[i] += re;
    }
#pragma unroll
    for (int i = 0; i < 32; ++i) {
        if (i >= N) break; // uniform
        const double r = block_sum<256>(acc[i], red);
        if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * N + i] = r;
    }
}

// partial[blk][sigma * N + i] = Re sum conj(a) b over elements whose species-sigma state has
// site i occupied: the gradient of an on-site potential eps_{i sigma} n_{i sigma}
template <bool CPLX>
__global__ __launch_bounds__(256) void k_site_dens(OtfArgs p, int N, const double *__restrict__ a,
                                                   const double *__restrict__ b, double *__restrict__ partial)
{
    __shared__ double red[4];
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t D = p.dim_up * p.dim_dn;
    double au[32], ad[32];
#pragma unroll
    for (int i = 0; i < 32; ++i) au[i] = ad[i] = 0.0;
    if (g < D) {
        const int64_t u = g / p.dim_dn, d = g - u * p.dim_dn;
        const uint32_t su = (uint32_t)p.state_up[u], sd = (uint32_t)p.state_dn[d];
        double re;
        if (CPLX) re = a[2 * g] * b[2 * g] + a[2 * g + 1] * b[2 * g + 1];
        else re = a[g] * b[g];
#pragma unroll
        for (int i = 0; i < 32; ++i) {
            if ((su >> i) & 1u) au[i] += re;
            if ((sd >> i) & 1u) ad[i] += re;
        }
    }
#pragma unroll
    for (int i = 0; i < 32; ++i) {
        if (i >= N) break; // uniform
        const double r = block_sum<256>(au[i], red);
        if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * 2 * N + i] = r;
        const double q = block_sum<256>(ad[i], red);
        if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * 2 * N + N + i] = q;
    }
}

// sum nrows x ncols partials over rows (fixed order) -> out[ncols]
__global__ __launch_bounds__(256) void k_colsum(const double *__restrict__ partial, int64_t nrows, int ncols,
                                                double *__restrict__ out)
{
    __shared__ double red[4];
    for (int c = 0; c < ncols; ++c) {
        double s = 0.0;
        for (int64_t r = threadIdx.x; r < nrows; r += 256) s += partial[r * ncols + c];
        s = block_sum<256>(s, red);
        if (threadIdx.x == 0) out[c] = s;
    }
}

} // namespace ed

using namespace ed;

// per-thread workspace of the bilinear / site-sum entry points
static thread_local ed::Arena t_adj_arena;

extern "C" int ed_bond_bilinear(ed_hubbard hh, int species, int complex_vectors, const void *d_a, const void *d_b,
                                double *out, void *stream)
{
    const Hubbard *h = reinterpret_cast<const Hubbard *>(hh);
    if (!h || !d_a || !d_b || !out || (species != 0 && species != 1))
        return fail(ED_EINVAL, "ed_bond_bilinear: bad arguments");
    if (!h->otf.up.ops_own || !h->otf.dn.ops_own)
        return fail(ED_EINVAL, "ed_bond_bilinear: operator needs ed_hubbard_create_ops");
    if (h->has_fermi)
        return fail(ED_EUNSUPPORTED, "ed_bond_bilinear: Insert_1Fop / Insert_2Fop calls are not supported");
    const OtfSpecies &S = species ? h->otf.dn : h->otf.up;
    hipStream_t st = stream_of(stream);
    const int64_t D = h->otf.dim_up * h->otf.dim_dn;
    const int64_t nb = (D + 255) / 256;
    Scratch SC(t_adj_arena, st);
    ED_TRY(SC.reserve(Scratch::round(sizeof(double) * (size_t)nb * 4 * kMaxBondOps) + Scratch::round(sizeof(double) * 4 * kMaxBondOps)));
    double *dpart = (double *)SC.take(sizeof(double) * (size_t)nb * 4 * kMaxBondOps), *dout = (double *)SC.take(sizeof(double) * 4 * kMaxBondOps);
    int s = ED_OK;
    std::vector<double> host(4 * kMaxBondOps);
    for (int o0 = 0; o0 < S.nops && !s; o0 += kMaxBondOps) {
        const int no = std::min(kMaxBondOps, S.nops - o0);
        const bool c = complex_vectors != 0;
        if (species == 0) {
            if (c) hipLaunchKernelGGL((k_bond_bilinear<true, false>), dim3((unsigned)nb), dim3(256), 0, st, h->otf, o0, no, (const double *)d_a, (const double *)d_b, dpart);
            else hipLaunchKernelGGL((k_bond_bilinear<false, false>), dim3((unsigned)nb), dim3(256), 0, st, h->otf, o0, no, (const double *)d_a, (const double *)d_b, dpart);
        } else {
            if (c) hipLaunchKernelGGL((k_bond_bilinear<true, true>), dim3((unsigned)nb), dim3(256), 0, st, h->otf, o0, no, (const double *)d_a, (const double *)d_b, dpart);
            else hipLaunchKernelGGL((k_bond_bilinear<false, true>), dim3((unsigned)nb), dim3(256), 0, st, h->otf, o0, no, (const double *)d_a, (const double *)d_b, dpart);
        }
        hipLaunchKernelGGL(k_colsum, dim3(1), dim3(256), 0, st, dpart, nb, 4 * no, dout);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(host.data(), dout, sizeof(double) * 4 * no, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            s = fail(ED_EHIP, "ed_bond_bilinear: kernel failed");
            break;
        }
        for (int k = 0; k < no; ++k)
            for (int q = 0; q < 4; ++q) out[(size_t)(o0 + k) * 4 + q] = host[(size_t)k * 4 + q];
    }
    return s;
}

extern "C" int ed_site_double_occupancy(ed_hubbard hh, int complex_vectors, const void *d_a, const void *d_b,
                                        double *out, void *stream)
{
    const Hubbard *h = reinterpret_cast<const Hubbard *>(hh);
    if (!h || !d_a || !d_b || !out) return fail(ED_EINVAL, "ed_site_double_occupancy: bad arguments");
    hipStream_t st = stream_of(stream);
    const int N = h->N;
    OtfArgs p = h->otf;
    p.dim_up = h->args.dim_up;
    p.dim_dn = h->args.dim_dn;
    p.state_up = h->d_state_up;
    p.state_dn = h->d_state_dn;
    const int64_t D = p.dim_up * p.dim_dn, nb = (D + 255) / 256;
    Scratch SC(t_adj_arena, st);
    ED_TRY(SC.reserve(Scratch::round(sizeof(double) * (size_t)nb * N) + Scratch::round(sizeof(double) * N)));
    double *dpart = (double *)SC.take(sizeof(double) * (size_t)nb * N), *dout = (double *)SC.take(sizeof(double) * N);
    int s = ED_OK;
    if (!s) {
        if (complex_vectors)
            hipLaunchKernelGGL(k_site_docc<true>, dim3((unsigned)nb), dim3(256), 0, st, p, N, (const double *)d_a, (const double *)d_b, dpart);
        else
            hipLaunchKernelGGL(k_site_docc<false>, dim3((unsigned)nb), dim3(256), 0, st, p, N, (const double *)d_a, (const double *)d_b, dpart);
        hipLaunchKernelGGL(k_colsum, dim3(1), dim3(256), 0, st, dpart, nb, N, dout);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(out, dout, sizeof(double) * N, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            s = fail(ED_EHIP, "ed_site_double_occupancy: kernel failed");
    }
    return s;
}

extern "C" int ed_site_density(ed_hubbard hh, int complex_vectors, const void *d_a, const void *d_b, double *out,
                               void *stream)
{
    const Hubbard *h = reinterpret_cast<const Hubbard *>(hh);
    if (!h || !d_a || !d_b || !out) return fail(ED_EINVAL, "ed_site_density: bad arguments");
    hipStream_t st = stream_of(stream);
    const int N = h->N;
    OtfArgs p = h->otf;
    p.dim_up = h->args.dim_up;
    p.dim_dn = h->args.dim_dn;
    p.state_up = h->d_state_up;
    p.state_dn = h->d_state_dn;
    const int64_t D = p.dim_up * p.dim_dn, nb = (D + 255) / 256;
    Scratch SC(t_adj_arena, st);
    ED_TRY(SC.reserve(Scratch::round(sizeof(double) * (size_t)nb * 2 * N) + Scratch::round(sizeof(double) * 2 * N)));
    double *dpart = (double *)SC.take(sizeof(double) * (size_t)nb * 2 * N), *dout = (double *)SC.take(sizeof(double) * 2 * N);
    int s = ED_OK;
    if (!s) {
        if (complex_vectors)
            hipLaunchKernelGGL(k_site_dens<true>, dim3((unsigned)nb), dim3(256), 0, st, p, N, (const double *)d_a, (const double *)d_b, dpart);
        else
            hipLaunchKernelGGL(k_site_dens<false>, dim3((unsigned)nb), dim3(256), 0, st, p, N, (const double *)d_a, (const double *)d_b, dpart);
        hipLaunchKernelGGL(k_colsum, dim3(1), dim3(256), 0, st, dpart, nb, 2 * N, dout);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(out, dout, sizeof(double) * 2 * N, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            s = fail(ED_EHIP, "ed_site_density: kernel failed");
    }
    return s;
}

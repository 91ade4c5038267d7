// ed_s2.hip -- total-spin operator y = (S+S- + S-S+)/2 x on the (N_up, N_dn) sector.
//
// Reference: TotalSpin::Hubbard + kernel::Hubbard_mv_ (TotalSpin/include/
// HubMatrix_TotalSpin.cuh:20-66, 126-173).  Per element (up, dn): for every site pair
// i < j whose spins are opposite and exchanged between i and j, gather x at the
// exchanged states (up^ex, dn^ex) with sign (-1)^{nb_up + nb_dn + 1}, nb = occupied
// sites strictly between i and j in the exchanged states; plus popcount(up^dn)/2 on
// the diagonal.  Gather form, one lane per output element, address tables (2^N
// entries per species) read through the L2.  The driver forms
// S(S+1) = (Q_up - Q_dn)^2/4 + <psi|y> (ConservedTotalSpin_Kagome_randomHop.cu:121-123).
#include "ed_common.hpp"

namespace ed {

struct S2 {
    int N;
    int64_t dim_up, dim_dn;
    int32_t *state_up, *state_dn, *addr_up, *addr_dn;
};

__device__ __forceinline__ int between(uint32_t s, int i, int j)
{
    // i < j: occupied sites strictly between
    const uint32_t m = ((1u << (j - i - 1)) - 1u) << (i + 1);
    return __popc(s & m);
}

template <bool CPLX>
__global__ __launch_bounds__(256) void k_s2(int N, int64_t dim_up, int64_t dim_dn, const int32_t *__restrict__ su,
                                            const int32_t *__restrict__ sd, const int32_t *__restrict__ au,
                                            const int32_t *__restrict__ ad, const double *__restrict__ x,
                                            double *__restrict__ y)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= dim_up * dim_dn) return;
    const int64_t sv = g / dim_dn, id = g - sv * dim_dn;
    const uint32_t up = (uint32_t)su[sv], dn = (uint32_t)sd[id];
    double re = 0.0, im = 0.0;
    const uint32_t diff = up ^ dn; // sites with opposite spins (or one empty)
    for (int i = 0; i < N; ++i) {
        const uint32_t iu = (up >> i) & 1u, idn = (dn >> i) & 1u;
        if (iu == idn) continue;
        for (int j = i + 1; j < N; ++j) {
            const uint32_t ju = (up >> j) & 1u, jd = (dn >> j) & 1u;
            if (iu != jd || idn != ju) continue;
            const uint32_t ex = (1u << i) | (1u << j);
            const uint32_t up_ = up ^ ex, dn_ = dn ^ ex;
            const int fs = between(up_, i, j) + between(dn_, i, j);
            const int64_t src = (int64_t)au[up_] * dim_dn + ad[dn_];
            const double sg = (fs & 1) ? 1.0 : -1.0;
            if (CPLX) {
                re += sg * x[2 * src];
                im += sg * x[2 * src + 1];
            } else {
                re += sg * x[src];
            }
        }
    }
    const double coef = 0.5 * (double)__popc(diff);
    if (CPLX) {
        y[2 * g] = re + coef * x[2 * g];
        y[2 * g + 1] = im + coef * x[2 * g + 1];
    } else {
        y[g] = re + coef * x[g];
    }
}

} // namespace ed

using namespace ed;

extern "C" int ed_s2_create(int N, int64_t dim_up, const int32_t *d_stateUP, const int32_t *d_addressUP,
                            int64_t dim_dn, const int32_t *d_stateDN, const int32_t *d_addressDN, ed_s2 *out)
{
    if (!out || N < 1 || N > 31 || dim_up < 1 || dim_dn < 1 || !d_stateUP || !d_addressUP || !d_stateDN ||
        !d_addressDN)
        return fail(ED_EINVAL, "ed_s2_create: bad arguments");
    *out = nullptr;
    S2 *s = new S2{N, dim_up, dim_dn, nullptr, nullptr, nullptr, nullptr};
    const size_t full = (size_t)1 << N;
    int st = dmalloc((void **)&s->state_up, sizeof(int32_t) * (size_t)dim_up);
    if (!st) st = dmalloc((void **)&s->state_dn, sizeof(int32_t) * (size_t)dim_dn);
    if (!st) st = dmalloc((void **)&s->addr_up, sizeof(int32_t) * full);
    if (!st) st = dmalloc((void **)&s->addr_dn, sizeof(int32_t) * full);
    if (!st && (hipMemcpy(s->state_up, d_stateUP, sizeof(int32_t) * dim_up, hipMemcpyDeviceToDevice) != hipSuccess ||
                hipMemcpy(s->state_dn, d_stateDN, sizeof(int32_t) * dim_dn, hipMemcpyDeviceToDevice) != hipSuccess ||
                hipMemcpy(s->addr_up, d_addressUP, sizeof(int32_t) * full, hipMemcpyDeviceToDevice) != hipSuccess ||
                hipMemcpy(s->addr_dn, d_addressDN, sizeof(int32_t) * full, hipMemcpyDeviceToDevice) != hipSuccess))
        st = fail(ED_EHIP, "ed_s2_create: copies failed");
    if (st) {
        dfree(s->state_up);
        dfree(s->state_dn);
        dfree(s->addr_up);
        dfree(s->addr_dn);
        delete s;
        return st;
    }
    *out = reinterpret_cast<ed_s2>(s);
    return ED_OK;
}

extern "C" int ed_s2_destroy(ed_s2 ss)
{
    S2 *s = reinterpret_cast<S2 *>(ss);
    if (!s) return ED_OK;
    dfree(s->state_up);
    dfree(s->state_dn);
    dfree(s->addr_up);
    dfree(s->addr_dn);
    delete s;
    return ED_OK;
}

extern "C" int ed_s2_apply(ed_s2 ss, int complex_vectors, const void *d_x, void *d_y, void *stream)
{
    const S2 *s = reinterpret_cast<const S2 *>(ss);
    if (!s || !d_x || !d_y || d_x == d_y) return fail(ED_EINVAL, "ed_s2_apply: bad arguments");
    const int64_t D = s->dim_up * s->dim_dn;
    const int64_t grid = (D + 255) / 256;
    if (grid > 0x7fffffffLL) return fail(ED_EUNSUPPORTED, "ed_s2_apply: grid too large");
    if (complex_vectors)
        hipLaunchKernelGGL(k_s2<true>, dim3((unsigned)grid), dim3(256), 0, stream_of(stream), s->N, s->dim_up,
                           s->dim_dn, s->state_up, s->state_dn, s->addr_up, s->addr_dn, (const double *)d_x,
                           (double *)d_y);
    else
        hipLaunchKernelGGL(k_s2<false>, dim3((unsigned)grid), dim3(256), 0, stream_of(stream), s->N, s->dim_up,
                           s->dim_dn, s->state_up, s->state_dn, s->addr_up, s->addr_dn, (const double *)d_x,
                           (double *)d_y);
    ED_CHECK_LAUNCH();
    return ED_OK;
}

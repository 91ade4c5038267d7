// ed_hubbard.hip -- the on-the-fly Hubbard H.v (Siro-Harju split) and the fused
// Lanczos kernels built on it.
//
// Reference: struct Hubbard + Hubbard_mv_ (MBL_test/include/HubMatrix_GPU.cuh:30-182).
//   y[u,d] = sum_s AxUP[s,u] x[AjUP[s,u], d]          (H_up (x) 1, rows mix)
//          + sum_s AxDN[s,d] x[u, AjDN[s,d]]          (1 (x) H_dn, columns mix)
//          + U popcount(up_u & dn_d & mask) x[u,d]    (diagonal)
// with psi row-major [u*dim_dn + d].  Per output element the terms are summed in the
// reference order: UP slots ascending, DN slots ascending, then U.  Slots that hold
// an exact 0.0 are dropped (the reference skips them for UP and adds 0*x for DN).
//
// Data layout in HBM (built once per operator from the reference's ELL tables):
//   UP side  -- CSR over rows u: the <= S nonzero slots of one row are shared by all
//               columns of that row, so a workgroup stages them in LDS and each lane
//               streams x[col*dim_dn + d] for consecutive d (coalesced 512 B per wave).
//   DN side  -- sliced ELL over columns d, slices of 64 columns (one wavefront), width
//               = the slice's longest row, entry k of column d at base + 64k + (d&63):
//               table loads are coalesced; padding entries point at the own column
//               with value 0.
//   diag     -- computed from the two state words (popcount, or a per-site U vector).
// All element offsets are 64-bit (D = 2.36e9 at 18 sites).
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "ed_common.hpp"
#include "ed_hubbard.hpp"

namespace ed {

// ------------------------------------------------------------------------------------
// compaction of the ELL tables
// ------------------------------------------------------------------------------------
template <bool CPLX>
__global__ void k_row_nnz(int64_t count, int S, const double *__restrict__ Ax, int32_t *__restrict__ nnz)
{
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= count) return;
    int c = 0;
    for (int s = 0; s < S; ++s) {
        const int64_t idx = (int64_t)s * count + a;
        const bool nz = CPLX ? (Ax[2 * idx] != 0.0 || Ax[2 * idx + 1] != 0.0) : (Ax[idx] != 0.0);
        c += nz ? 1 : 0;
    }
    nnz[a] = c;
}

// CSR fill (UP side)
template <bool CPLX>
__global__ void k_fill_csr(int64_t count, int S, const double *__restrict__ Ax, const int32_t *__restrict__ Aj,
                           const int64_t *__restrict__ ptr, int32_t *__restrict__ col, double *__restrict__ val)
{
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= count) return;
    int64_t o = ptr[a];
    for (int s = 0; s < S; ++s) {
        const int64_t idx = (int64_t)s * count + a;
        if (CPLX) {
            const double re = Ax[2 * idx], im = Ax[2 * idx + 1];
            if (re == 0.0 && im == 0.0) continue;
            val[2 * o] = re;
            val[2 * o + 1] = im;
        } else {
            const double v = Ax[idx];
            if (v == 0.0) continue;
            val[o] = v;
        }
        col[o] = Aj[idx];
        ++o;
    }
}

// sliced-ELL fill (DN side): slice = 64 consecutive columns
template <bool CPLX>
__global__ void k_fill_sell(int64_t count, int S, const double *__restrict__ Ax, const int32_t *__restrict__ Aj,
                            const int64_t *__restrict__ sptr, int32_t *__restrict__ col, double *__restrict__ val)
{
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= count) return;
    const int64_t sl = d >> 6;
    const int lane = (int)(d & 63);
    const int64_t base = sptr[sl];
    const int w = (int)((sptr[sl + 1] - base) >> 6);
    int k = 0;
    for (int s = 0; s < S; ++s) {
        const int64_t idx = (int64_t)s * count + d;
        double re, im = 0.0;
        if (CPLX) {
            re = Ax[2 * idx];
            im = Ax[2 * idx + 1];
            if (re == 0.0 && im == 0.0) continue;
        } else {
            re = Ax[idx];
            if (re == 0.0) continue;
        }
        const int64_t e = base + (int64_t)k * 64 + lane;
        col[e] = Aj[idx];
        if (CPLX) {
            val[2 * e] = re;
            val[2 * e + 1] = im;
        } else {
            val[e] = re;
        }
        ++k;
    }
    for (; k < w; ++k) {
        const int64_t e = base + (int64_t)k * 64 + lane;
        col[e] = (int32_t)d;
        if (CPLX) {
            val[2 * e] = 0.0;
            val[2 * e + 1] = 0.0;
        } else {
            val[e] = 0.0;
        }
    }
}

static int compact_species(const SpeciesIn &in, bool cplx, Species &out, bool as_csr, hipStream_t st)
{
    out.dim = in.dim;
    out.state = nullptr;
    const int64_t n = in.dim;
    const int B = 256;
    const unsigned grid = (unsigned)((n + B - 1) / B);
    int32_t *dnnz = nullptr;
    ED_TRY(dmalloc((void **)&dnnz, sizeof(int32_t) * (size_t)n));
    if (cplx) hipLaunchKernelGGL(k_row_nnz<true>, dim3(grid), dim3(B), 0, st, n, in.S, in.Ax, dnnz);
    else hipLaunchKernelGGL(k_row_nnz<false>, dim3(grid), dim3(B), 0, st, n, in.S, in.Ax, dnnz);
    std::vector<int32_t> nnz((size_t)n);
    if (hipMemcpyAsync(nnz.data(), dnnz, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        dfree(dnnz);
        return fail(ED_EHIP, "compact_species: nnz count failed");
    }
    dfree(dnnz);
    const size_t vsz = cplx ? 2 * sizeof(double) : sizeof(double);
    int maxw = 0;
    for (int32_t c : nnz) maxw = std::max(maxw, (int)c);
    out.max_nnz = maxw;
    if (as_csr) {
        std::vector<int64_t> ptr((size_t)n + 1);
        ptr[0] = 0;
        for (int64_t a = 0; a < n; ++a) ptr[a + 1] = ptr[a] + nnz[a];
        const int64_t tot = ptr[n];
        out.nent = tot;
        ED_TRY(dmalloc((void **)&out.ptr, sizeof(int64_t) * (size_t)(n + 1)));
        ED_TRY(dmalloc((void **)&out.col, sizeof(int32_t) * (size_t)std::max<int64_t>(tot, 1)));
        ED_TRY(dmalloc((void **)&out.val, vsz * (size_t)std::max<int64_t>(tot, 1)));
        ED_HIP(hipMemcpyAsync(out.ptr, ptr.data(), sizeof(int64_t) * (size_t)(n + 1), hipMemcpyHostToDevice, st));
        if (cplx)
            hipLaunchKernelGGL(k_fill_csr<true>, dim3(grid), dim3(B), 0, st, n, in.S, in.Ax, in.Aj, out.ptr,
                               out.col, out.val);
        else
            hipLaunchKernelGGL(k_fill_csr<false>, dim3(grid), dim3(B), 0, st, n, in.S, in.Ax, in.Aj, out.ptr,
                               out.col, out.val);
        ED_CHECK_LAUNCH();
        ED_HIP(hipStreamSynchronize(st)); // ptr (host vector) must outlive the copy
    } else {
        const int64_t nsl = (n + 63) / 64;
        std::vector<int64_t> sptr((size_t)nsl + 1);
        sptr[0] = 0;
        for (int64_t s = 0; s < nsl; ++s) {
            int w = 0;
            for (int64_t d = s * 64; d < std::min<int64_t>(n, s * 64 + 64); ++d) w = std::max(w, (int)nnz[d]);
            sptr[s + 1] = sptr[s] + (int64_t)w * 64;
        }
        const int64_t tot = sptr[nsl];
        out.nent = tot;
        ED_TRY(dmalloc((void **)&out.ptr, sizeof(int64_t) * (size_t)(nsl + 1)));
        ED_TRY(dmalloc((void **)&out.col, sizeof(int32_t) * (size_t)std::max<int64_t>(tot, 1)));
        ED_TRY(dmalloc((void **)&out.val, vsz * (size_t)std::max<int64_t>(tot, 1)));
        ED_HIP(hipMemcpyAsync(out.ptr, sptr.data(), sizeof(int64_t) * (size_t)(nsl + 1), hipMemcpyHostToDevice,
                              st));
        if (cplx)
            hipLaunchKernelGGL(k_fill_sell<true>, dim3(grid), dim3(B), 0, st, n, in.S, in.Ax, in.Aj, out.ptr,
                               out.col, out.val);
        else
            hipLaunchKernelGGL(k_fill_sell<false>, dim3(grid), dim3(B), 0, st, n, in.S, in.Ax, in.Aj, out.ptr,
                               out.col, out.val);
        ED_CHECK_LAUNCH();
        ED_HIP(hipStreamSynchronize(st));
    }
    return ED_OK;
}

// ------------------------------------------------------------------------------------
// value helpers (real / complex)
// ------------------------------------------------------------------------------------
template <typename V>
struct VT;
template <>
struct VT<double> {
    static __device__ __forceinline__ double zero() { return 0.0; }
    static __device__ __forceinline__ double scale(double s, double x) { return s * x; }
    static __device__ __forceinline__ double fma_rr(double a, double x, double acc) { return fma(a, x, acc); }
    static __device__ __forceinline__ double dotre(double a, double b) { return a * b; }
    static __device__ __forceinline__ double abs2(double a) { return a * a; }
};
template <>
struct VT<cplx> {
    static __device__ __forceinline__ cplx zero() { return cplx{0.0, 0.0}; }
    static __device__ __forceinline__ cplx scale(double s, cplx x) { return cplx{s * x.re, s * x.im}; }
    static __device__ __forceinline__ cplx fma_rr(double a, cplx x, cplx acc)
    {
        return cplx{fma(a, x.re, acc.re), fma(a, x.im, acc.im)};
    }
    static __device__ __forceinline__ double dotre(cplx a, cplx b) { return a.re * b.re + a.im * b.im; }
    static __device__ __forceinline__ double abs2(cplx a) { return a.re * a.re + a.im * a.im; }
};

// acc += a * x for table value type TV and vector type V
__device__ __forceinline__ double mac(double a, double x, double acc) { return fma(a, x, acc); }
__device__ __forceinline__ cplx mac(double a, cplx x, cplx acc)
{
    return cplx{fma(a, x.re, acc.re), fma(a, x.im, acc.im)};
}
__device__ __forceinline__ cplx mac(cplx a, cplx x, cplx acc)
{
    // (a.re + i a.im)(x.re + i x.im)
    return cplx{fma(a.re, x.re, fma(-a.im, x.im, acc.re)), fma(a.re, x.im, fma(a.im, x.re, acc.im))};
}

template <typename V>
__device__ __forceinline__ V ldv(const V *p, int64_t i)
{
    return p[i];
}

template <typename V>
__device__ __forceinline__ V vscale(double s, V x)
{
    return VT<V>::scale(s, x);
}

__device__ __forceinline__ double diag_value(int32_t su, int32_t sd, double U, int64_t mask,
                                             const double *__restrict__ Uvec)
{
    const uint32_t dbl = (uint32_t)(su & sd) & (uint32_t)mask;
    if (!Uvec) return U * (double)__popc(dbl);
    double s = 0.0;
    uint32_t m = dbl;
    while (m) {
        const int i = __ffs(m) - 1;
        s += Uvec[i];
        m &= m - 1;
    }
    return s;
}

// ------------------------------------------------------------------------------------
// H.v kernel.  One workgroup = one row u and a run of CPB = 256*CPT columns.
// FLAGS: HV_SCALE (x gathered as s*x, s = 1/b_dev[step]), HV_WRITEV (v[i] = s*x[i]),
//        HV_AXPY (y = Hx + c*w, c = -b_dev[step]), HV_DOT (block partial Re<v,y>).
// ------------------------------------------------------------------------------------

template <typename V, typename TV, int CPT, int FLAGS>
__global__ __launch_bounds__(256) void k_hv(HvArgs p, const V *__restrict__ x, V *y,
                                            V *__restrict__ vout, const V *w, const double *__restrict__ bdev,
                                            int step, double *__restrict__ partial)
{
    constexpr int B = 256;
    constexpr int CPB = B * CPT;
    constexpr int MAXUP = 512;
    __shared__ int32_t s_col[MAXUP];
    __shared__ TV s_val[MAXUP];
    __shared__ double s_red[B / 64];

    const int64_t u = blockIdx.x / p.blocks_per_row;
    const int64_t cb = blockIdx.x - u * p.blocks_per_row;
    const int64_t dim_dn = p.dim_dn;
    const int64_t row0 = u * dim_dn;

    const int64_t up0 = p.up_ptr[u];
    const int nup = (int)(p.up_ptr[u + 1] - up0);
    for (int t = threadIdx.x; t < nup; t += B) {
        s_col[t] = p.up_col[up0 + t];
        s_val[t] = reinterpret_cast<const TV *>(p.up_val)[up0 + t];
    }
    __syncthreads();

    double s = 1.0, c = 0.0;
    if (FLAGS & HV_SCALE) s = 1.0 / bdev[step];
    if (FLAGS & HV_AXPY) c = -bdev[step];
    const int32_t su = p.state_up[u];

    V acc[CPT];
    int64_t dcol[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        acc[k] = VT<V>::zero();
        dcol[k] = cb * CPB + (int64_t)k * B + threadIdx.x;
    }

    // UP term: rows col_j of x, same columns (coalesced streams)
    for (int j = 0; j < nup; ++j) {
        const int64_t rb = (int64_t)s_col[j] * dim_dn;
        const TV a = s_val[j];
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            if (dcol[k] < dim_dn) {
                V xv = x[rb + dcol[k]];
                if (FLAGS & HV_SCALE) xv = vscale(s, xv);
                acc[k] = mac(a, xv, acc[k]);
            }
        }
    }

    double dot = 0.0;
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int64_t d = dcol[k];
        if (d >= dim_dn) continue;
        // DN term: gathers inside row u
        const int64_t sl = d >> 6;
        const int lane = (int)(d & 63);
        const int64_t base = p.dn_sptr[sl];
        const int wdt = (int)((p.dn_sptr[sl + 1] - base) >> 6);
        const TV *dv = reinterpret_cast<const TV *>(p.dn_val);
        for (int e = 0; e < wdt; ++e) {
            const int64_t ei = base + (int64_t)e * 64 + lane;
            V xv = x[row0 + p.dn_col[ei]];
            if (FLAGS & HV_SCALE) xv = vscale(s, xv);
            acc[k] = mac(dv[ei], xv, acc[k]);
        }
        // diagonal
        V xd = x[row0 + d];
        if (FLAGS & HV_SCALE) xd = vscale(s, xd);
        if (p.has_diag) {
            const double dg = diag_value(su, p.state_dn[d], p.U, p.mask, p.Uvec);
            acc[k] = VT<V>::fma_rr(dg, xd, acc[k]);
        }
        V out = acc[k];
        if (FLAGS & HV_AXPY) out = VT<V>::fma_rr(c, w[row0 + d], out);
        y[row0 + d] = out;
        if (FLAGS & HV_WRITEV) vout[row0 + d] = xd;
        if (FLAGS & HV_DOT) dot += VT<V>::dotre(xd, out);
    }
    if (FLAGS & HV_DOT) {
        const double r = block_sum<B>(dot, s_red);
        if (threadIdx.x == 0) partial[blockIdx.x] = r;
    }
}

template <typename V, typename TV, int FLAGS>
static int launch_hv_cpt(const HvArgs &a0, int cpt, const V *x, V *y, V *vout, const V *w, const double *bdev,
                         int step, double *partial, hipStream_t st, int64_t *nblocks_out)
{
    HvArgs a = a0;
    const int64_t cpb = 256 * (int64_t)cpt;
    a.blocks_per_row = (a.dim_dn + cpb - 1) / cpb;
    const int64_t nb = a.blocks_per_row * a.dim_up;
    if (nb > 0x7fffffffLL) return fail(ED_EUNSUPPORTED, "H.v grid too large (%lld blocks)", (long long)nb);
    if (nblocks_out) *nblocks_out = nb;
    dim3 g((unsigned)nb), b(256);
    switch (cpt) {
    case 1: hipLaunchKernelGGL((k_hv<V, TV, 1, FLAGS>), g, b, 0, st, a, x, y, vout, w, bdev, step, partial); break;
    case 2: hipLaunchKernelGGL((k_hv<V, TV, 2, FLAGS>), g, b, 0, st, a, x, y, vout, w, bdev, step, partial); break;
    default: hipLaunchKernelGGL((k_hv<V, TV, 4, FLAGS>), g, b, 0, st, a, x, y, vout, w, bdev, step, partial); break;
    }
    ED_CHECK_LAUNCH();
    return ED_OK;
}

int choose_cpt(int64_t dim_up, int64_t dim_dn)
{
    // enough workgroups to cover 256 CUs several times over
    for (int cpt : {4, 2}) {
        const int64_t bpr = (dim_dn + 256 * cpt - 1) / (256 * cpt);
        if (bpr * dim_up >= 8192 && 256 * cpt <= dim_dn * 2) return cpt;
    }
    return 1;
}

// ------------------------------------------------------------------------------------
// Two-pass H.v (default).  Workgroups are dispatched round-robin over the 8 XCDs
// (workgroup b runs on XCD b % 8), so a physical id is remapped to a logical tile
// t = (b % 8) * per_xcd + b / 8: every XCD walks one contiguous range of tiles and its
// private 4 MB L2 holds that range's working set.
//   pass 1 (UP term): tiles = column slabs of W doubles x RB = 256/W rows, slab-major.
//            An XCD sweeps all rows of its slabs; the slab X[:, slab] (dim_up*W*8 bytes,
//            3.1 MB at W = 8 for 18 sites) stays in its L2, so the ~19 neighbour-row
//            reads per element hit L2 instead of HBM.  Writes the UP partial sum.
//   pass 2 (DN term + diagonal + Lanczos epilogue): tiles = R rows x 256 columns,
//            row-panel-major, so the rows being gathered stay in the XCD's L2 and each
//            DN table entry is loaded once for R rows.  Starts from the pass-1 partial,
//            so each element is summed in the reference order (UP, DN, then U).
// ------------------------------------------------------------------------------------
__device__ __forceinline__ int64_t xcd_tile(int64_t per_xcd)
{
    const int64_t b = blockIdx.x;
    return (b & 7) * per_xcd + (b >> 3);
}

template <typename V, typename TV, int W, bool SCALE>
__global__ __launch_bounds__(256) void k_hv_up(HvArgs p, int64_t ntiles, int64_t per_xcd, int64_t nrp, int maxnnz,
                                               const V *__restrict__ x, V *out, const double *__restrict__ bdev,
                                               int step)
{
    constexpr int RB = 256 / W;
    extern __shared__ double smem_d[];
    TV *s_val = reinterpret_cast<TV *>(smem_d);
    int32_t *s_col = reinterpret_cast<int32_t *>(s_val + RB * maxnnz);
    int32_t *s_n = s_col + RB * maxnnz;
    const int64_t t = xcd_tile(per_xcd);
    if (t >= ntiles) return; // uniform over the workgroup
    const int64_t slab = t / nrp, rp = t - slab * nrp;
    const int64_t u0 = rp * RB;
    const int tid = threadIdx.x;
    if (tid < RB) {
        const int64_t u = u0 + tid;
        s_n[tid] = (u < p.dim_up) ? (int)(p.up_ptr[u + 1] - p.up_ptr[u]) : 0;
    }
    for (int e = tid; e < RB * maxnnz; e += 256) {
        const int r = e / maxnnz, j = e - r * maxnnz;
        const int64_t u = u0 + r;
        if (u < p.dim_up) {
            const int64_t b0 = p.up_ptr[u];
            if (j < (int)(p.up_ptr[u + 1] - b0)) {
                s_col[e] = p.up_col[b0 + j];
                s_val[e] = reinterpret_cast<const TV *>(p.up_val)[b0 + j];
            }
        }
    }
    __syncthreads();
    const int r = tid / W, c = tid - r * W;
    const int64_t u = u0 + r, d = slab * W + c;
    if (u >= p.dim_up || d >= p.dim_dn) return;
    double s = 1.0;
    if (SCALE) s = 1.0 / bdev[step];
    V acc = VT<V>::zero();
    const int n = s_n[r];
    const int32_t *cl = s_col + r * maxnnz;
    const TV *vl = s_val + r * maxnnz;
    for (int j = 0; j < n; ++j) {
        V xv = x[(int64_t)cl[j] * p.dim_dn + d];
        if (SCALE) xv = vscale(s, xv);
        acc = mac(vl[j], xv, acc);
    }
    out[u * p.dim_dn + d] = acc;
}

template <typename V, typename TV, int R, int FLAGS>
__global__ __launch_bounds__(256) void k_hv_dn(HvArgs p, int64_t ntiles, int64_t per_xcd, int64_t ncb, bool has_up,
                                               const V *__restrict__ x, const V *ypart, V *y, V *vout, const V *w,
                                               const double *__restrict__ bdev, int step,
                                               double *__restrict__ partial)
{
    __shared__ double s_red[4];
    const int64_t t = xcd_tile(per_xcd);
    if (t >= ntiles) return; // uniform over the workgroup
    const int64_t rp = t / ncb, cb = t - rp * ncb;
    const int64_t u0 = rp * R, d = cb * 256 + threadIdx.x;
    const int64_t dim_dn = p.dim_dn;
    double s = 1.0, c = 0.0;
    if (FLAGS & HV_SCALE) s = 1.0 / bdev[step];
    if (FLAGS & HV_AXPY) c = -bdev[step];
    double dot = 0.0;
    if (d < dim_dn) {
        V acc[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int64_t u = u0 + r;
            acc[r] = (has_up && u < p.dim_up) ? ypart[u * dim_dn + d] : VT<V>::zero();
        }
        const int64_t sl = d >> 6;
        const int lane = (int)(d & 63);
        const int64_t base = p.dn_sptr[sl];
        const int wdt = (int)((p.dn_sptr[sl + 1] - base) >> 6);
        const TV *dv = reinterpret_cast<const TV *>(p.dn_val);
        for (int e = 0; e < wdt; ++e) {
            const int64_t ei = base + (int64_t)e * 64 + lane;
            const int64_t col = p.dn_col[ei];
            const TV a = dv[ei];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int64_t u = u0 + r;
                if (u < p.dim_up) {
                    V xv = x[u * dim_dn + col];
                    if (FLAGS & HV_SCALE) xv = vscale(s, xv);
                    acc[r] = mac(a, xv, acc[r]);
                }
            }
        }
        const int32_t sd = p.state_dn[d];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int64_t u = u0 + r;
            if (u >= p.dim_up) continue;
            const int64_t i = u * dim_dn + d;
            V xd = x[i];
            if (FLAGS & HV_SCALE) xd = vscale(s, xd);
            if (p.has_diag) acc[r] = VT<V>::fma_rr(diag_value(p.state_up[u], sd, p.U, p.mask, p.Uvec), xd, acc[r]);
            V o = acc[r];
            if (FLAGS & HV_AXPY) o = VT<V>::fma_rr(c, w[i], o);
            y[i] = o;
            if (FLAGS & HV_WRITEV) vout[i] = xd;
            if (FLAGS & HV_DOT) dot += VT<V>::dotre(xd, o);
        }
    }
    if (FLAGS & HV_DOT) {
        const double rsum = block_sum<256>(dot, s_red);
        if (threadIdx.x == 0) partial[t] = rsum;
    }
}

struct Plan2 {
    int W, R;
    int64_t nslab, nrp1, tiles1, per1;
    int64_t ncb, nrp2, tiles2, per2;
};

static Plan2 plan2(const Hubbard &h)
{
    Plan2 P;
    const int64_t du = h.args.dim_up, dd = h.args.dim_dn;
    P.W = (h.hv_w == 16) ? 16 : 8;
    const int RB = 256 / P.W;
    P.nslab = (dd + P.W - 1) / P.W;
    P.nrp1 = (du + RB - 1) / RB;
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

template <typename V, typename TV, int FLAGS>
static int launch_hv2(const Hubbard &h, const V *x, V *y, V *vout, const V *w, const double *bdev, int step,
                      double *partial, hipStream_t st, int64_t *nblocks_out)
{
    const Plan2 P = plan2(h);
    if (P.per1 * 8 > 0x7fffffffLL || P.per2 * 8 > 0x7fffffffLL) return fail(ED_EUNSUPPORTED, "H.v grid too large");
    const int maxnnz = std::max(1, h.up.max_nnz);
    const bool has_up = h.up.nent > 0;
    // Lanczos: the UP partial goes into vout (pass 2 reads it back before writing v there),
    // because y aliases w (v_{N-1}) which pass 2 still needs.
    V *upbuf = (FLAGS & HV_WRITEV) ? vout : y;
    constexpr bool SC = (FLAGS & HV_SCALE) != 0;
    if (has_up) {
        const int RB = 256 / P.W;
        const size_t sh = (size_t)RB * maxnnz * (sizeof(TV) + sizeof(int32_t)) + RB * sizeof(int32_t);
        if (P.W == 16)
            hipLaunchKernelGGL((k_hv_up<V, TV, 16, SC>), dim3((unsigned)(P.per1 * 8)), dim3(256), sh, st, h.args,
                               P.tiles1, P.per1, P.nrp1, maxnnz, x, upbuf, bdev, step);
        else
            hipLaunchKernelGGL((k_hv_up<V, TV, 8, SC>), dim3((unsigned)(P.per1 * 8)), dim3(256), sh, st, h.args,
                               P.tiles1, P.per1, P.nrp1, maxnnz, x, upbuf, bdev, step);
        ED_CHECK_LAUNCH();
    }
    dim3 g2((unsigned)(P.per2 * 8)), b2(256);
    switch (P.R) {
    case 8:
        hipLaunchKernelGGL((k_hv_dn<V, TV, 8, FLAGS>), g2, b2, 0, st, h.args, P.tiles2, P.per2, P.ncb, has_up, x,
                           upbuf, y, vout, w, bdev, step, partial);
        break;
    case 4:
        hipLaunchKernelGGL((k_hv_dn<V, TV, 4, FLAGS>), g2, b2, 0, st, h.args, P.tiles2, P.per2, P.ncb, has_up, x,
                           upbuf, y, vout, w, bdev, step, partial);
        break;
    case 2:
        hipLaunchKernelGGL((k_hv_dn<V, TV, 2, FLAGS>), g2, b2, 0, st, h.args, P.tiles2, P.per2, P.ncb, has_up, x,
                           upbuf, y, vout, w, bdev, step, partial);
        break;
    default:
        hipLaunchKernelGGL((k_hv_dn<V, TV, 1, FLAGS>), g2, b2, 0, st, h.args, P.tiles2, P.per2, P.ncb, has_up, x,
                           upbuf, y, vout, w, bdev, step, partial);
        break;
    }
    ED_CHECK_LAUNCH();
    if (nblocks_out) *nblocks_out = P.tiles2;
    return ED_OK;
}

template <int FLAGS>
int hv_dispatch(const Hubbard &h, bool cvec, const void *x, void *y, void *vout, const void *w, const double *bdev,
                int step, double *partial, hipStream_t st, int64_t *nblocks)
{
    if (!cvec && h.cplx_tables) return fail(ED_EINVAL, "complex tables need complex vectors");
    if (h.algo == 3) return otf_dispatch<FLAGS>(h, cvec, x, y, vout, w, bdev, step, partial, st, nblocks);
    if (h.algo == 1) {
        const int cpt = choose_cpt(h.args.dim_up, h.args.dim_dn);
        if (!cvec)
            return launch_hv_cpt<double, double, FLAGS>(h.args, cpt, (const double *)x, (double *)y,
                                                        (double *)vout, (const double *)w, bdev, step, partial, st,
                                                        nblocks);
        if (h.cplx_tables)
            return launch_hv_cpt<cplx, cplx, FLAGS>(h.args, cpt, (const cplx *)x, (cplx *)y, (cplx *)vout,
                                                    (const cplx *)w, bdev, step, partial, st, nblocks);
        return launch_hv_cpt<cplx, double, FLAGS>(h.args, cpt, (const cplx *)x, (cplx *)y, (cplx *)vout,
                                                  (const cplx *)w, bdev, step, partial, st, nblocks);
    }
    if (!cvec)
        return launch_hv2<double, double, FLAGS>(h, (const double *)x, (double *)y, (double *)vout,
                                                 (const double *)w, bdev, step, partial, st, nblocks);
    if (h.cplx_tables)
        return launch_hv2<cplx, cplx, FLAGS>(h, (const cplx *)x, (cplx *)y, (cplx *)vout, (const cplx *)w, bdev,
                                             step, partial, st, nblocks);
    return launch_hv2<cplx, double, FLAGS>(h, (const cplx *)x, (cplx *)y, (cplx *)vout, (const cplx *)w, bdev, step,
                                           partial, st, nblocks);
}

int64_t hv_num_blocks(const Hubbard &h)
{
    if (h.algo == 3) return otf_num_blocks(h);
    if (h.algo == 1) {
        const int cpt = choose_cpt(h.args.dim_up, h.args.dim_dn);
        const int64_t cpb = 256 * (int64_t)cpt;
        return ((h.args.dim_dn + cpb - 1) / cpb) * h.args.dim_up;
    }
    return plan2(h).tiles2;
}

int hv_plain(const Hubbard &h, bool cvec, const void *x, void *y, hipStream_t st)
{
    return hv_dispatch<0>(h, cvec, x, y, nullptr, nullptr, nullptr, 0, nullptr, st, nullptr);
}

// Lanczos step, first half: v[i] = x[i]/b[step]; y = H v - b[step] w (step > 0); partial Re<v,y>
int hv_lanczos(const Hubbard &h, bool cvec, const void *f, void *y, void *v, const void *w, const double *bdev,
               int step, double *partial, hipStream_t st, int64_t *nblocks)
{
    if (step == 0)
        return hv_dispatch<HV_SCALE | HV_WRITEV | HV_DOT>(h, cvec, f, y, v, nullptr, bdev, step, partial, st,
                                                          nblocks);
    return hv_dispatch<HV_SCALE | HV_WRITEV | HV_AXPY | HV_DOT>(h, cvec, f, y, v, w, bdev, step, partial, st,
                                                                 nblocks);
}

// Ritz rebuild: same recurrence without the dot
int hv_lanczos_nodot(const Hubbard &h, bool cvec, const void *f, void *y, void *v, const void *w,
                     const double *bdev, int step, hipStream_t st)
{
    if (step == 0)
        return hv_dispatch<HV_SCALE | HV_WRITEV>(h, cvec, f, y, v, nullptr, bdev, step, nullptr, st, nullptr);
    return hv_dispatch<HV_SCALE | HV_WRITEV | HV_AXPY>(h, cvec, f, y, v, w, bdev, step, nullptr, st, nullptr);
}

void Hubbard::release()
{
    for (Species *s : {&up, &dn}) {
        dfree(s->ptr);
        dfree(s->col);
        dfree(s->val);
        s->ptr = nullptr;
        s->col = nullptr;
        s->val = nullptr;
    }
    dfree(d_state_up);
    dfree(d_state_dn);
    dfree(d_uvec);
    dfree(d_addr_up);
    dfree(d_addr_dn);
    dfree(d_up_order);
    dfree(d_up_tab);
    dfree(d_up_cnt);
    d_up_order = nullptr;
    d_up_tab = nullptr;
    d_up_cnt = nullptr;
    dfree(otf.up.ops_own);
    dfree(otf.dn.ops_own);
    slab_free(slab);
    slab = nullptr;
    d_state_up = d_state_dn = nullptr;
    d_uvec = nullptr;
    d_addr_up = d_addr_dn = nullptr;
    otf.up.ops_own = otf.dn.ops_own = nullptr;
}

} // namespace ed

using namespace ed;

extern "C" int ed_hubbard_create(int N, int64_t dim_up, int S_up, const double *d_AxUP, const int32_t *d_AjUP,
                                 const int32_t *d_stateUP, int64_t dim_dn, int S_dn, const double *d_AxDN,
                                 const int32_t *d_AjDN, const int32_t *d_stateDN, int complex_tables, double U,
                                 int64_t mask, const double *Uvec, ed_hubbard *out)
{
    if (!out || N < 1 || N > 31 || dim_up < 1 || dim_dn < 1 || S_up < 1 || S_dn < 1 || !d_AxUP || !d_AjUP ||
        !d_AxDN || !d_AjDN || !d_stateUP || !d_stateDN)
        return fail(ED_EINVAL, "ed_hubbard_create: bad arguments");
    if (dim_dn > 0x7fffffffLL || dim_up > 0x7fffffffLL)
        return fail(ED_EUNSUPPORTED, "ed_hubbard_create: species dimension beyond int32");
    *out = nullptr;
    Hubbard *h = new Hubbard();
    h->N = N;
    h->cplx_tables = complex_tables != 0;
    hipStream_t st = nullptr;
    SpeciesIn iu{dim_up, S_up, d_AxUP, d_AjUP}, idn{dim_dn, S_dn, d_AxDN, d_AjDN};
    int s = compact_species(iu, h->cplx_tables, h->up, true, st);
    if (s == ED_OK) s = compact_species(idn, h->cplx_tables, h->dn, false, st);
    if (s == ED_OK) s = dmalloc((void **)&h->d_state_up, sizeof(int32_t) * (size_t)dim_up);
    if (s == ED_OK) s = dmalloc((void **)&h->d_state_dn, sizeof(int32_t) * (size_t)dim_dn);
    if (s == ED_OK && Uvec) {
        s = dmalloc((void **)&h->d_uvec, sizeof(double) * 32);
        if (s == ED_OK) {
            double tmp[32] = {0};
            for (int i = 0; i < N; ++i) tmp[i] = Uvec[i];
            h->h_uvec.assign(tmp, tmp + N);
            if (hipMemcpy(h->d_uvec, tmp, sizeof tmp, hipMemcpyHostToDevice) != hipSuccess)
                s = fail(ED_EHIP, "ed_hubbard_create: Uvec upload failed");
        }
    }
    if (s == ED_OK) {
        if (hipMemcpy(h->d_state_up, d_stateUP, sizeof(int32_t) * (size_t)dim_up, hipMemcpyDeviceToDevice) !=
                hipSuccess ||
            hipMemcpy(h->d_state_dn, d_stateDN, sizeof(int32_t) * (size_t)dim_dn, hipMemcpyDeviceToDevice) !=
                hipSuccess)
            s = fail(ED_EHIP, "ed_hubbard_create: state copy failed");
    }
    if (s != ED_OK) {
        h->release();
        delete h;
        return s;
    }
    h->U = U;
    h->mask = (mask < 0) ? (int64_t)(((uint64_t)1 << N) - 1) : mask;
    HvArgs &a = h->args;
    a.dim_up = dim_up;
    a.dim_dn = dim_dn;
    a.up_ptr = h->up.ptr;
    a.up_col = h->up.col;
    a.up_val = h->up.val;
    a.dn_sptr = h->dn.ptr;
    a.dn_col = h->dn.col;
    a.dn_val = h->dn.val;
    a.state_up = h->d_state_up;
    a.state_dn = h->d_state_dn;
    a.U = U;
    a.mask = h->mask;
    a.Uvec = h->d_uvec;
    a.has_diag = (U != 0.0 || h->d_uvec) ? 1 : 0;
    a.blocks_per_row = 1;
    // kernel variant: ED_HV_ALGO=1 selects the single-pass row kernel (for A/B timing),
    // ED_HV_W (8|16) the pass-1 slab width, ED_HV_RMAX the largest pass-2 row panel
    if (const char *e = getenv("ED_HV_ALGO")) h->algo = atoi(e) == 1 ? 1 : 2;
    if (const char *e = getenv("ED_HV_W")) h->hv_w = atoi(e);
    if (const char *e = getenv("ED_HV_RMAX")) h->hv_rmax = atoi(e);
    if (const char *e = getenv("ED_HV_DNMASK")) h->hv_dnmask = atoi(e) != 0;
    if (const char *e = getenv("ED_HV_DN2")) h->hv_dn2 = atoi(e) != 0;
    if (const char *e = getenv("ED_HV_DNK")) h->hv_dnk = atoi(e);
    if (const char *e = getenv("ED_HV_UPXCD")) h->hv_upxcd = atoi(e) != 0;
    if (const char *e = getenv("ED_HV_UPTAB")) h->hv_uptab = atoi(e) != 0;
    // pass 1 keeps 256/W rows of UP lists in LDS: rows with very many slots use the
    // single-pass kernel instead (its LDS holds one row's list)
    if ((size_t)32 * h->up.max_nnz * (2 * sizeof(double) + sizeof(int32_t)) + 256 > 65536) h->algo = 1;
    if (h->up.max_nnz > 512) {
        h->release();
        delete h;
        return fail(ED_EUNSUPPORTED, "ed_hubbard_create: a row has %d > 512 nonzero slots", h->up.max_nnz);
    }
    *out = reinterpret_cast<ed_hubbard>(h);
    return ED_OK;
}

extern "C" int ed_hubbard_destroy(ed_hubbard hh)
{
    Hubbard *h = reinterpret_cast<Hubbard *>(hh);
    if (!h) return ED_OK;
    h->release();
    delete h;
    return ED_OK;
}

extern "C" int ed_hubbard_info(ed_hubbard hh, int64_t *dim_up, int64_t *dim_dn, int *complex_tables)
{
    const Hubbard *h = reinterpret_cast<const Hubbard *>(hh);
    if (!h) return fail(ED_EINVAL, "ed_hubbard_info: null handle");
    if (dim_up) *dim_up = h->args.dim_up;
    if (dim_dn) *dim_dn = h->args.dim_dn;
    if (complex_tables) *complex_tables = h->cplx_tables ? 1 : 0;
    return ED_OK;
}

extern "C" int ed_hubbard_apply(ed_hubbard hh, int complex_vectors, const void *d_x, void *d_y, void *stream)
{
    const Hubbard *h = reinterpret_cast<const Hubbard *>(hh);
    if (!h || !d_x || !d_y) return fail(ED_EINVAL, "ed_hubbard_apply: null argument");
    if (d_x == d_y) return fail(ED_EINVAL, "ed_hubbard_apply: x and y must differ");
    return hv_plain(*h, complex_vectors != 0, d_x, d_y, stream_of(stream));
}

// Reverse Cuthill-McKee order of the hopping graph of one sector: vertices are the colex
// ranks, edges the hopping calls (a hop links s and s ^ (1<<i | 1<<j) when exactly one of the
// two sites is occupied).  Pass 1 of the on-the-fly H.v walks its rows in this order, so the
// rows read for one row were mostly read for rows visited shortly before and are still in L2
// (breadth-first level sets instead of colex ranks, whose hops on high sites jump far).
int ed::up_row_order(int N, int Q, const ed_op *ops, int nops, std::vector<int32_t> &order)
{
    int64_t dim = 0;
    ED_TRY(ed_basis_count(N, Q, &dim));
    if (dim > (int64_t)1 << 30) return fail(ED_EUNSUPPORTED, "row order: sector too large");
    int64_t C[33][33] = {};
    for (int n = 0; n <= 32; ++n) {
        C[n][0] = 1;
        for (int k = 1; k <= n; ++k) C[n][k] = C[n - 1][k - 1] + (k <= n - 1 ? C[n - 1][k] : 0);
    }
    auto rank = [&](uint32_t s) {
        int64_t r = 0;
        for (int k = 1; s; ++k, s &= s - 1) r += C[__builtin_ctz(s)][k];
        return r;
    };
    std::vector<uint32_t> st((size_t)dim);
    uint64_t v = Q ? (((uint64_t)1 << Q) - 1) : 0;
    for (int64_t r = 0; r < dim; ++r) {
        st[(size_t)r] = (uint32_t)v;
        if (r + 1 < dim) { // next integer with Q bits set
            const uint64_t c = v & (~v + 1), t = v + c;
            v = (((t ^ v) >> 2) / c) | t;
        }
    }
    std::vector<uint32_t> hm;
    for (int o = 0; o < nops; ++o)
        if (ops[o].kind != ED_OP_DENSITY && ops[o].kind != ED_OP_DENSITY2 && ops[o].i != ops[o].j)
            hm.push_back((1u << ops[o].i) | (1u << ops[o].j));
    auto for_nbrs = [&](int64_t r, auto &&f) {
        const uint32_t s = st[(size_t)r];
        for (uint32_t m : hm) {
            const uint32_t b = s & m;
            if (b && b != m) f(rank(s ^ m));
        }
    };
    std::vector<int32_t> deg((size_t)dim);
    for (int64_t r = 0; r < dim; ++r) {
        int n = 0;
        for_nbrs(r, [&](int64_t) { ++n; });
        deg[(size_t)r] = n;
    }
    std::vector<int32_t> starts((size_t)dim);
    for (int64_t r = 0; r < dim; ++r) starts[(size_t)r] = (int32_t)r;
    std::stable_sort(starts.begin(), starts.end(), [&](int32_t a, int32_t b) { return deg[a] < deg[b]; });
    std::vector<char> seen((size_t)dim, 0);
    order.clear();
    order.reserve((size_t)dim);
    std::vector<int32_t> nb;
    for (int32_t s0 : starts) {
        if (seen[s0]) continue;
        seen[s0] = 1;
        size_t head = order.size();
        order.push_back(s0);
        while (head < order.size()) {
            nb.clear();
            for_nbrs(order[head++], [&](int64_t q) {
                if (!seen[q]) {
                    seen[q] = 1;
                    nb.push_back((int32_t)q);
                }
            });
            std::sort(nb.begin(), nb.end(), [&](int32_t a, int32_t b) { return deg[a] != deg[b] ? deg[a] < deg[b] : a < b; });
            order.insert(order.end(), nb.begin(), nb.end());
        }
    }
    std::reverse(order.begin(), order.end());
    return ED_OK;
}

// Tile-shape autotuner for the on-the-fly H.v: times the candidates on the caller's vectors
// and keeps the fastest in the handle.  Pass 1 (slab width, row order) and pass 2 (kernel, row
// panels) are separate kernels whose times add, so the search runs in two stages: the pass-1
// candidates with the default pass 2, then the pass-2 candidates with the best pass 1.  All
// candidates add the same terms in the same order per element, so the result of H.v does not
// depend on the choice.
extern "C" int ed_hubbard_autotune(ed_hubbard hh, int complex_vectors, const void *d_x, void *d_y, int reps,
                                   ed_hv_tuning *best, void *stream)
{
    Hubbard *h = reinterpret_cast<Hubbard *>(hh);
    if (!h || !d_x || !d_y || d_x == d_y || reps < 1) return fail(ED_EINVAL, "ed_hubbard_autotune: bad arguments");
    if (h->algo != 3) return fail(ED_EUNSUPPORTED, "ed_hubbard_autotune: needs an operator from ed_hubbard_create_ops");
    hipStream_t st = stream_of(stream);
    struct Cand {
        int w, ord, tab, dn2, dnk, rmax;
    };
    const Hubbard saved = *h;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ED_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(ED_EHIP, "ed_hubbard_autotune: event creation failed");
    }
    auto set = [&](const Cand &c) {
        h->hv_w = c.w;
        h->hv_uporder = c.ord;
        h->hv_uptab = c.tab;
        h->hv_dn2 = c.dn2;
        h->hv_dnk = c.dnk ? c.dnk : saved.hv_dnk;
        h->hv_rmax = c.rmax;
        h->hv_dnmask = 1;
    };
    auto timeit = [&](const Cand &c, float &ms) {
        set(c);
        int s = hv_plain(*h, complex_vectors != 0, d_x, d_y, st); // warm-up
        if (!s && hipEventRecord(e0, st) != hipSuccess) s = fail(ED_EHIP, "ed_hubbard_autotune: event record failed");
        for (int r = 0; r < reps && !s; ++r) s = hv_plain(*h, complex_vectors != 0, d_x, d_y, st);
        if (!s && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                   hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            s = fail(ED_EHIP, "ed_hubbard_autotune: timing failed");
        return s;
    };
    Cand bc{128, h->d_up_order ? 1 : 0, h->d_up_tab ? 1 : 0, 1, 2, 8};
    float best_ms = 0.0f;
    int s = ED_OK;
    bool first = true;
    // stage 1: pass-1 shapes
    for (int w : {32, 64, 128, 256})
        for (int ord = h->d_up_order ? 1 : 0; ord >= 0; --ord)
            for (int tab = h->d_up_tab ? 1 : 0; tab >= 0 && !s; --tab) {
                const Cand c{w, ord, tab, 1, 2, 8};
                float ms = 0.0f;
                if ((s = timeit(c, ms))) break;
                if (first || ms < best_ms) {
                    best_ms = ms;
                    bc = c;
                    first = false;
                }
            }
    // stage 2: pass-2 shapes with the chosen pass 1
    const Cand dns[] = {{0, 0, 0, 1, 4, 8}, {0, 0, 0, 1, 2, 4}, {0, 0, 0, 1, 1, 8}, {0, 0, 0, 0, 0, 8},
                        {0, 0, 0, 0, 0, 4}};
    const Cand b1 = bc;
    for (const Cand &d : dns) {
        if (s) break;
        const Cand c{b1.w, b1.ord, b1.tab, d.dn2, d.dnk, d.rmax};
        float ms = 0.0f;
        if ((s = timeit(c, ms))) break;
        if (ms < best_ms) {
            best_ms = ms;
            bc = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    h->hv_w = saved.hv_w;
    h->hv_uporder = saved.hv_uporder;
    h->hv_uptab = saved.hv_uptab;
    h->hv_dn2 = saved.hv_dn2;
    h->hv_dnk = saved.hv_dnk;
    h->hv_rmax = saved.hv_rmax;
    h->hv_dnmask = saved.hv_dnmask;
    if (s) return s;
    set(bc);
    if (best) {
        best->slab_width = bc.w;
        best->row_order = bc.ord;
        best->up_table = bc.tab;
        best->dn_cached = bc.dn2;
        best->dn_panels = bc.dn2 ? h->hv_dnk : 1;
        best->dn_rows = bc.rmax;
        best->ms_per_apply = best_ms / reps;
    }
    return ED_OK;
}

extern "C" int ed_hubbard_create_ops(int N, int Q_up, const ed_op *ops_up, int nops_up, int Q_dn, const ed_op *ops_dn,
                                     int nops_dn, int complex_tables, double U, int64_t mask, const double *Uvec,
                                     ed_hubbard *out)
{
    if (!out || N < 1 || N > 31 || Q_up < 0 || Q_dn < 0 || Q_up > N || Q_dn > N || nops_up < 0 || nops_dn < 0 ||
        (nops_up && !ops_up) || (nops_dn && !ops_dn))
        return fail(ED_EINVAL, "ed_hubbard_create_ops: bad arguments");
    *out = nullptr;
    int64_t cu, cd;
    ED_TRY(ed_basis_count(N, Q_up, &cu));
    ED_TRY(ed_basis_count(N, Q_dn, &cd));
    const int Su = kZSlots * Q_up + 1, Sd = kZSlots * Q_dn + 1;
    const size_t full = (size_t)1 << N;
    const size_t vs = complex_tables ? 2 : 1;
    int32_t *su = nullptr, *au = nullptr, *sd = nullptr, *ad = nullptr, *ju = nullptr, *jd = nullptr;
    double *xu = nullptr, *xd = nullptr;
    ed_hubbard hh = nullptr;
    int s = ED_OK;
    auto cleanup = [&] {
        dfree(su);
        dfree(sd);
        dfree(ju);
        dfree(jd);
        dfree(xu);
        dfree(xd);
        dfree(au);
        dfree(ad);
    };
    do {
        if ((s = dmalloc((void **)&su, sizeof(int32_t) * cu)) || (s = dmalloc((void **)&sd, sizeof(int32_t) * cd)) ||
            (s = dmalloc((void **)&au, sizeof(int32_t) * full)) ||
            (s = dmalloc((void **)&ad, sizeof(int32_t) * full)) ||
            (s = dmalloc((void **)&xu, sizeof(double) * vs * cu * Su)) ||
            (s = dmalloc((void **)&xd, sizeof(double) * vs * cd * Sd)) ||
            (s = dmalloc((void **)&ju, sizeof(int32_t) * cu * Su)) ||
            (s = dmalloc((void **)&jd, sizeof(int32_t) * cd * Sd)))
            break;
        if ((s = ed_basis_build(N, Q_up, su, au, nullptr)) || (s = ed_basis_build(N, Q_dn, sd, ad, nullptr))) break;
        if ((s = ed_tables_build(N, Q_up, su, au, cu, ops_up, nops_up, complex_tables, xu, ju, nullptr)) ||
            (s = ed_tables_build(N, Q_dn, sd, ad, cd, ops_dn, nops_dn, complex_tables, xd, jd, nullptr)))
            break;
        if ((s = ed_hubbard_create(N, cu, Su, xu, ju, su, cd, Sd, xd, jd, sd, complex_tables, U, mask, Uvec, &hh)))
            break;
        Hubbard *h = reinterpret_cast<Hubbard *>(hh);
        h->d_addr_up = au;
        h->d_addr_dn = ad;
        au = ad = nullptr; // owned by h now
        OtfArgs &o = h->otf;
        o.dim_up = cu;
        o.dim_dn = cd;
        o.state_up = h->d_state_up;
        o.state_dn = h->d_state_dn;
        o.U = h->U;
        o.mask = h->mask;
        o.Uvec = h->d_uvec;
        o.has_diag = h->args.has_diag;
        if ((s = otf_prepare(o.up, ops_up, nops_up, h->d_addr_up, nullptr)) ||
            (s = otf_prepare(o.dn, ops_dn, nops_dn, h->d_addr_dn, nullptr)))
            break;
        h->h_ops_up.resize((size_t)nops_up);
        h->h_ops_dn.resize((size_t)nops_dn);
        if ((nops_up && hipMemcpy(h->h_ops_up.data(), o.up.ops, sizeof(OtfOp) * nops_up, hipMemcpyDeviceToHost) !=
                            hipSuccess) ||
            (nops_dn && hipMemcpy(h->h_ops_dn.data(), o.dn.ops, sizeof(OtfOp) * nops_dn, hipMemcpyDeviceToHost) !=
                            hipSuccess)) {
            s = fail(ED_EHIP, "ed_hubbard_create_ops: operator readback failed");
            break;
        }
        for (int o = 0; o < nops_up; ++o) h->has_fermi |= ops_up[o].kind == ED_OP_FERMI1 || ops_up[o].kind == ED_OP_FERMI2;
        for (int o = 0; o < nops_dn; ++o) h->has_fermi |= ops_dn[o].kind == ED_OP_FERMI1 || ops_dn[o].kind == ED_OP_FERMI2;
        int algo = (nops_up <= 64 && nops_dn <= 64) ? 3 : h->algo;
        if (const char *e = getenv("ED_HV_ALGO")) algo = atoi(e);
        // the on-the-fly hop generation covers the particle-conserving calls only
        if (algo == 3 && (nops_up > 64 || nops_dn > 64 || h->has_fermi)) algo = 2;
        if (algo < 1 || algo > 3) algo = 2;
        h->algo = algo;
        // on-the-fly pass 1 reads 128-column slabs (one 1 KB row segment per wavefront load,
        // 16 B per lane); narrower slabs re-walk every row of x once per slab (18-site H.v, UP
        // pass: W=8 82 ms, W=16 ~55 ms, W=64 37 ms, W=128 ~33 ms)
        if (algo == 3 && !getenv("ED_HV_W")) h->hv_w = 128;
        // pass-1 row order for vectors far beyond L2 + MALL (1 GiB of FP64 and up);
        // ED_HV_UPORDER=1 builds it at any size, 0 never
        const char *eo = getenv("ED_HV_UPORDER");
        const bool want = eo ? atoi(eo) != 0 : (double)cu * (double)cd >= (double)(1 << 27);
        if (algo == 3 && want && cu > 1) {
            std::vector<int32_t> ord;
            if ((s = up_row_order(N, Q_up, ops_up, nops_up, ord))) break;
            if ((s = dmalloc((void **)&h->d_up_order, sizeof(int32_t) * ord.size()))) break;
            if (hipMemcpy(h->d_up_order, ord.data(), sizeof(int32_t) * ord.size(), hipMemcpyHostToDevice) !=
                hipSuccess) {
                s = fail(ED_EHIP, "ed_hubbard_create_ops: row order upload failed");
                break;
            }
        }
        if (algo == 3 && (s = otf_build_up_table(*h, nullptr))) break;
    } while (0);
    cleanup();
    if (s) {
        if (hh) ed_hubbard_destroy(hh);
        return s;
    }
    *out = hh;
    return ED_OK;
}

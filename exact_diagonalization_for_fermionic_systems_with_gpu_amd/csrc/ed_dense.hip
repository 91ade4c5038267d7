// ed_dense.hip -- the dense full-spectrum path: the Hubbard and H_TBq matrices written out as
// dense D x D arrays, their full eigendecomposition (rocSOLVER syevd / heevd) and the
// eigenbasis propagation exp(-iH dt) of tED.
//
// Reference:
//   Hubbard::Write / write_      (MBL_test/include/HubMatrix_GPU.cuh:184-230)
//   Hubbard::initialize           (HubMatrix_GPU.cuh:250-270)
//   H_TBq::Write                  (MBL_test/include/fermi_operator.h:227-235)
//   cusolverDsyevd / Zheevd       (MBL_test/include/EIGEN_cusolver.h:16-63, 65-112)
//   tED<C>::Make_eigen, tED_time_evol (HubMatrix_GPU.cuh:730-797)
//
// Layouts (the reference's): H[row*D + col] row-major, row = u*dim_dn + d.  The eigensolver
// returns eigenvalues ascending and eigenvector k contiguous at A[k*n ...] (a column of the
// column-major result), which is how tED_time_evol reads them (HubMatrix_GPU.cuh:789).
//
// The dense matrix is O(D^2): these entry points serve the small sectors of the MBL quench and
// TotalSpinFull drivers (D up to a few 10^4), not the Lanczos-scale path.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cmath>
#include <vector>

#include "ed_common.hpp"
#include "ed_thread.hpp"
#include "ed_hubbard.hpp"

namespace ed {

namespace {

template <typename TV>
struct DV;
template <>
struct DV<double> {
    static __device__ __forceinline__ bool nz(double a) { return a != 0.0; }
    static __device__ __forceinline__ void add(double *H, int64_t i, double a) { H[i] += a; }
};
template <>
struct DV<cplx> {
    static __device__ __forceinline__ bool nz(cplx a) { return a.re != 0.0 || a.im != 0.0; }
    static __device__ __forceinline__ void add(double *H, int64_t i, cplx a)
    {
        H[2 * i] += a.re;
        H[2 * i + 1] += a.im;
    }
};

__device__ __forceinline__ double dense_diag(int32_t su, int32_t sd, const HvArgs &p)
{
    const uint32_t dbl = (uint32_t)(su & sd) & (uint32_t)p.mask;
    if (!p.Uvec) return p.U * (double)__popc(dbl);
    double s = 0.0;
    uint32_t m = dbl;
    while (m) {
        const int i = __ffs(m) - 1;
        s += p.Uvec[i];
        m &= m - 1;
    }
    return s;
}

// write_ (HubMatrix_GPU.cuh:184-217): one thread per row r = u*dim_dn + d of H; the nonzero UP
// slots (in slot order), then the nonzero DN slots, then the interaction, each added into
// H[r*D + column].  The compact tables keep exactly the nonzero slots in slot order (DN padding
// entries carry the value 0 and are skipped like the reference's empty slots).
template <typename TV>
__global__ __launch_bounds__(256) void k_write_dense(HvArgs p, double *__restrict__ H)
{
    const int64_t dd = p.dim_dn, D = p.dim_up * dd;
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= D) return;
    const int64_t u = r / dd, d = r - u * dd;
    const int64_t row = r * D;
    const TV *uv = reinterpret_cast<const TV *>(p.up_val);
    for (int64_t e = p.up_ptr[u]; e < p.up_ptr[u + 1]; ++e) {
        const TV a = uv[e];
        if (DV<TV>::nz(a)) DV<TV>::add(H, row + (int64_t)p.up_col[e] * dd + d, a);
    }
    const int64_t sl = d >> 6;
    const int lane = (int)(d & 63);
    const int64_t base = p.dn_sptr[sl];
    const int w = (int)((p.dn_sptr[sl + 1] - base) >> 6);
    const TV *dv = reinterpret_cast<const TV *>(p.dn_val);
    for (int k = 0; k < w; ++k) {
        const int64_t e = base + (int64_t)k * 64 + lane;
        const TV a = dv[e];
        if (DV<TV>::nz(a)) DV<TV>::add(H, row + u * dd + p.dn_col[e], a);
    }
    if (p.has_diag) H[(sizeof(TV) / sizeof(double)) * (row + r)] += dense_diag(p.state_up[u], p.state_dn[d], p);
}

// H_TBq::Write (fermi_operator.h:227-235): slot-major walk, M[j*count + Aj] += Ax for every
// nonzero slot; one thread per row j visits its slots in ascending order, which is the order in
// which the reference's loop reaches the entries of row j.
template <typename TV>
__global__ __launch_bounds__(256) void k_tables_write_dense(int64_t count, int S, const double *__restrict__ Ax,
                                                            const int32_t *__restrict__ Aj, double *__restrict__ M)
{
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= count) return;
    const TV *ax = reinterpret_cast<const TV *>(Ax);
    for (int s = 0; s < S; ++s) {
        const int64_t idx = (int64_t)s * count + j;
        const TV a = ax[idx];
        if (DV<TV>::nz(a)) DV<TV>::add(M, j * count + Aj[idx], a);
    }
}

__global__ __launch_bounds__(256) void k_zero(int64_t n, double *__restrict__ a)
{
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) a[i] = 0.0;
}

__global__ __launch_bounds__(256) void k_conj(int64_t n, double *__restrict__ a)
{
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        a[2 * i + 1] = -a[2 * i + 1];
}

// tED_time_evol, step 1 (HubMatrix_GPU.cuh:788-793): c_k = e^{-i W_k dt} <V_k|phi0>, one workgroup
// per eigenvector k (a contiguous column), fixed-order block reduction.
template <typename TV>
__global__ __launch_bounds__(256) void k_ted_project(int64_t n, const double *__restrict__ V,
                                                     const double *__restrict__ W, const cplx *__restrict__ phi0,
                                                     double dt, cplx *__restrict__ c)
{
    __shared__ double s_red[4];
    const int64_t k = blockIdx.x;
    const TV *col = reinterpret_cast<const TV *>(V) + k * n;
    double re = 0.0, im = 0.0;
    for (int64_t j = threadIdx.x; j < n; j += 256) {
        const cplx p = phi0[j];
        if constexpr (sizeof(TV) == sizeof(double)) {
            const double v = reinterpret_cast<const double *>(col)[j];
            re = fma(v, p.re, re);
            im = fma(v, p.im, im);
        } else {
            const cplx v = reinterpret_cast<const cplx *>(col)[j];
            // conj(v) p
            re = fma(v.re, p.re, fma(v.im, p.im, re));
            im = fma(v.re, p.im, fma(-v.im, p.re, im));
        }
    }
    re = block_sum<256>(re, s_red);
    im = block_sum<256>(im, s_red);
    if (threadIdx.x == 0) {
        double s, cs;
        sincos(W[k] * dt, &s, &cs);
        // (cos - i sin)(re + i im)
        c[k] = cplx{cs * re + s * im, cs * im - s * re};
    }
}

// step 2 (HubMatrix_GPU.cuh:794): partial[kc][j] = sum over the KC eigenvectors of chunk kc of
// c_k V_k[j]; lanes walk consecutive j (coalesced column reads), k ascending within the chunk.
constexpr int kTedKC = 64;
template <typename TV>
__global__ __launch_bounds__(256) void k_ted_expand(int64_t n, const double *__restrict__ V,
                                                    const cplx *__restrict__ c, cplx *__restrict__ partial)
{
    __shared__ cplx s_c[kTedKC];
    const int64_t kc = blockIdx.y;
    const int64_t k0 = kc * kTedKC;
    const int kn = (int)((n - k0) < kTedKC ? (n - k0) : kTedKC);
    if (threadIdx.x < kn) s_c[threadIdx.x] = c[k0 + threadIdx.x];
    __syncthreads();
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    double re = 0.0, im = 0.0;
    const TV *base = reinterpret_cast<const TV *>(V) + k0 * n + j;
    for (int q = 0; q < kn; ++q) {
        const cplx a = s_c[q];
        if constexpr (sizeof(TV) == sizeof(double)) {
            const double v = reinterpret_cast<const double *>(base)[(int64_t)q * n];
            re = fma(a.re, v, re);
            im = fma(a.im, v, im);
        } else {
            const cplx v = reinterpret_cast<const cplx *>(base)[(int64_t)q * n];
            re = fma(a.re, v.re, fma(-a.im, v.im, re));
            im = fma(a.re, v.im, fma(a.im, v.re, im));
        }
    }
    partial[kc * n + j] = cplx{re, im};
}

// step 3: phi[j] += sum over chunks in ascending order
__global__ __launch_bounds__(256) void k_ted_reduce(int64_t n, int64_t nchunk, const cplx *__restrict__ partial,
                                                    cplx *__restrict__ phi)
{
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    cplx acc = phi[j];
    for (int64_t kc = 0; kc < nchunk; ++kc) {
        const cplx p = partial[kc * n + j];
        acc.re += p.re;
        acc.im += p.im;
    }
    phi[j] = acc;
}

unsigned grid_for(int64_t n)
{
    const int64_t g = (n + 255) / 256;
    return (unsigned)(g < 65536 ? (g > 0 ? g : 1) : 65536);
}

rocblas_handle dense_handle() { return thread_blas_handle(); }

} // namespace

} // namespace ed

using namespace ed;

extern "C" int ed_dense_zero(int64_t n, int complex_values, void *d_A, void *stream)
{
    if (n < 0 || (n > 0 && !d_A)) return fail(ED_EINVAL, "ed_dense_zero: bad arguments");
    if (n == 0) return ED_OK;
    const int64_t m = n * n * (complex_values ? 2 : 1);
    hipLaunchKernelGGL(k_zero, dim3(grid_for(m)), dim3(256), 0, stream_of(stream), m, (double *)d_A);
    ED_CHECK_LAUNCH();
    return ED_OK;
}

extern "C" int ed_hubbard_write_dense(ed_hubbard hh, void *d_H, void *stream)
{
    if (!hh || !d_H) return fail(ED_EINVAL, "ed_hubbard_write_dense: bad arguments");
    const Hubbard *h = reinterpret_cast<const Hubbard *>(hh);
    const int64_t D = h->args.dim_up * h->args.dim_dn;
    if (D > 65536) return fail(ED_EINVAL, "ed_hubbard_write_dense: D = %lld is too large for a dense matrix", (long long)D);
    hipStream_t st = stream_of(stream);
    const unsigned g = (unsigned)((D + 255) / 256);
    if (h->cplx_tables) hipLaunchKernelGGL(k_write_dense<cplx>, dim3(g), dim3(256), 0, st, h->args, (double *)d_H);
    else hipLaunchKernelGGL(k_write_dense<double>, dim3(g), dim3(256), 0, st, h->args, (double *)d_H);
    ED_CHECK_LAUNCH();
    return ED_OK;
}

extern "C" int ed_tables_write_dense(int64_t count, int S, int complex_tables, const double *d_Ax,
                                     const int32_t *d_Aj, void *d_M, void *stream)
{
    if (count < 1 || count > 65536 || S < 1 || !d_Ax || !d_Aj || !d_M)
        return fail(ED_EINVAL, "ed_tables_write_dense: bad arguments");
    hipStream_t st = stream_of(stream);
    const unsigned g = (unsigned)((count + 255) / 256);
    if (complex_tables)
        hipLaunchKernelGGL(k_tables_write_dense<cplx>, dim3(g), dim3(256), 0, st, count, S, d_Ax, d_Aj, (double *)d_M);
    else
        hipLaunchKernelGGL(k_tables_write_dense<double>, dim3(g), dim3(256), 0, st, count, S, d_Ax, d_Aj,
                           (double *)d_M);
    ED_CHECK_LAUNCH();
    return ED_OK;
}

extern "C" int ed_dense_eigh(int64_t n, int complex_values, void *d_A, double *d_W, void *stream)
{
    if (n < 1 || n > 46340 || !d_A || !d_W) return fail(ED_EINVAL, "ed_dense_eigh: bad arguments (n = %lld)", (long long)n);
    hipStream_t st = stream_of(stream);
    rocblas_handle h = dense_handle();
    if (!h) return fail(ED_EHIP, "ed_dense_eigh: rocblas handle creation failed");
    rocblas_set_stream(h, st);
    static thread_local Arena t_dense_arena;
    Scratch S(t_dense_arena, st);
    const size_t bE = sizeof(double) * (size_t)n, bI = sizeof(rocblas_int);
    ED_TRY(S.reserve(Scratch::round(bE) + Scratch::round(bI)));
    double *dE = (double *)S.take(bE);
    rocblas_int *dinfo = (rocblas_int *)S.take(bI);
    rocblas_status rs;
    // the row-major H read as a column-major array is H^T = conj(H): real symmetric matrices are
    // unchanged, and for complex ones the eigenvectors come out conjugated and are conjugated back
    if (complex_values)
        rs = rocsolver_zheevd(h, rocblas_evect_original, rocblas_fill_lower, (rocblas_int)n,
                              (rocblas_double_complex *)d_A, (rocblas_int)n, d_W, dE, dinfo);
    else
        rs = rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_lower, (rocblas_int)n, (double *)d_A,
                              (rocblas_int)n, d_W, dE, dinfo);
    if (rs != rocblas_status_success) return fail(ED_EHIP, "ed_dense_eigh: rocsolver syevd/heevd failed (%d)", (int)rs);
    if (complex_values) {
        hipLaunchKernelGGL(k_conj, dim3(grid_for(n * n)), dim3(256), 0, st, n * n, (double *)d_A);
        ED_CHECK_LAUNCH();
    }
    rocblas_int info = 0;
    ED_HIP(hipMemcpyAsync(&info, dinfo, sizeof info, hipMemcpyDeviceToHost, st));
    ED_HIP(hipStreamSynchronize(st));
    if (info != 0) return fail(ED_ENOCONV, "ed_dense_eigh: syevd/heevd info = %d", (int)info);
    return ED_OK;
}

extern "C" int ed_ted_time_evol(int64_t n, int complex_values, const void *d_V, const double *d_W,
                                const void *d_phi0, double dt, void *d_phi, void *stream)
{
    if (n < 1 || n > 46340 || !d_V || !d_W || !d_phi0 || !d_phi || d_phi == d_phi0)
        return fail(ED_EINVAL, "ed_ted_time_evol: bad arguments");
    hipStream_t st = stream_of(stream);
    const int64_t nchunk = (n + kTedKC - 1) / kTedKC;
    static thread_local Arena t_ted_arena;
    Scratch S(t_ted_arena, st);
    const size_t bC = sizeof(cplx) * (size_t)n, bP = sizeof(cplx) * (size_t)(nchunk * n);
    ED_TRY(S.reserve(Scratch::round(bC) + Scratch::round(bP)));
    cplx *c = (cplx *)S.take(bC);
    cplx *part = (cplx *)S.take(bP);
    const double *V = (const double *)d_V;
    const unsigned gj = (unsigned)((n + 255) / 256);
    if (complex_values) {
        hipLaunchKernelGGL(k_ted_project<cplx>, dim3((unsigned)n), dim3(256), 0, st, n, V, d_W, (const cplx *)d_phi0,
                           dt, c);
        hipLaunchKernelGGL(k_ted_expand<cplx>, dim3(gj, (unsigned)nchunk), dim3(256), 0, st, n, V, c, part);
    } else {
        hipLaunchKernelGGL(k_ted_project<double>, dim3((unsigned)n), dim3(256), 0, st, n, V, d_W,
                           (const cplx *)d_phi0, dt, c);
        hipLaunchKernelGGL(k_ted_expand<double>, dim3(gj, (unsigned)nchunk), dim3(256), 0, st, n, V, c, part);
    }
    hipLaunchKernelGGL(k_ted_reduce, dim3(gj), dim3(256), 0, st, n, nchunk, part, (cplx *)d_phi);
    ED_CHECK_LAUNCH();
    return ED_OK;
}

// ed_blas.hip -- vector BLAS-1 for the Lanczos path (replaces Thrust_Lin,
// MBL_test/include/HubMatrix_GPU.cuh:291-471) and the seed-235 start vector
// (GenRand, HubMatrix_GPU.cuh:504-539).
//
// Reductions are deterministic: every block reduces a fixed 2048-element chunk in a
// fixed tree, writes one partial, and a single-block finalize sums the partials in
// index order.  Plain and fused paths therefore give bit-identical scalars, and a
// run gives the same scalars every time.  Complex vectors are handled as 2n reals
// (Re<a,b> and |a|^2 are real dots of the interleaved pairs).
#include <mutex>

#include "ed_common.hpp"
#include "ed_thread.hpp"
#include "ed_hubbard.hpp"

namespace ed {

constexpr int kBB = 256;             // block size of the BLAS kernels
constexpr int kChunk = kBB * 8;      // elements per block

int64_t blas_num_blocks(int64_t nreal) { return (nreal + kChunk - 1) / kChunk; }

// mode 0: dst = sum; 1: dst = sqrt(sum)
__global__ __launch_bounds__(1024) void k_finalize(const double *__restrict__ partial, int64_t n, double *dst,
                                                   int mode)
{
    __shared__ double red[16];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 1024) s += partial[i];
    s = block_sum<1024>(s, red);
    if (threadIdx.x == 0) *dst = (mode == 1) ? sqrt(s) : s;
}

int finalize_sum(const double *partial, int64_t n, double *dst, int mode, hipStream_t st)
{
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, st, partial, n, dst, mode);
    ED_CHECK_LAUNCH();
    return ED_OK;
}

__global__ __launch_bounds__(kBB) void k_dot(int64_t n, const double *__restrict__ a, const double *__restrict__ b,
                                             double *__restrict__ partial)
{
    __shared__ double red[kBB / 64];
    const int64_t base = (int64_t)blockIdx.x * kChunk + threadIdx.x;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kChunk / kBB; ++k) {
        const int64_t i = base + (int64_t)k * kBB;
        if (i < n) s = fma(a[i], b[i], s);
    }
    s = block_sum<kBB>(s, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// imaginary part of <a,b> for interleaved complex: sum a.re*b.im - a.im*b.re
__global__ __launch_bounds__(kBB) void k_dot_im(int64_t ncplx, const double *__restrict__ a,
                                                const double *__restrict__ b, double *__restrict__ partial)
{
    __shared__ double red[kBB / 64];
    const int64_t base = (int64_t)blockIdx.x * kChunk + threadIdx.x;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kChunk / kBB; ++k) {
        const int64_t i = base + (int64_t)k * kBB;
        if (i < ncplx) s += a[2 * i] * b[2 * i + 1] - a[2 * i + 1] * b[2 * i];
    }
    s = block_sum<kBB>(s, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

int dot_partials(int64_t nreal, const double *a, const double *b, double *partial, int64_t *nblocks, hipStream_t st)
{
    const int64_t nb = blas_num_blocks(nreal);
    if (nblocks) *nblocks = nb;
    hipLaunchKernelGGL(k_dot, dim3((unsigned)nb), dim3(kBB), 0, st, nreal, a, b, partial);
    ED_CHECK_LAUNCH();
    return ED_OK;
}

// f = -a[step]*v + f ; partial sum f^2
__global__ __launch_bounds__(kBB) void k_lanczos_update(int64_t n, const double *__restrict__ v, double *__restrict__ f,
                                                        const double *__restrict__ adev, int step,
                                                        double *__restrict__ partial)
{
    __shared__ double red[kBB / 64];
    const double c = -adev[step];
    const int64_t base = (int64_t)blockIdx.x * kChunk + threadIdx.x;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kChunk / kBB; ++k) {
        const int64_t i = base + (int64_t)k * kBB;
        if (i < n) {
            const double r = fma(c, v[i], f[i]);
            f[i] = r;
            s = fma(r, r, s);
        }
    }
    s = block_sum<kBB>(s, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

int lanczos_update(int64_t nreal, const double *v, double *f, const double *adev, int step, double *partial,
                   int64_t *nblocks, hipStream_t st)
{
    const int64_t nb = blas_num_blocks(nreal);
    if (nblocks) *nblocks = nb;
    hipLaunchKernelGGL(k_lanczos_update, dim3((unsigned)nb), dim3(kBB), 0, st, nreal, v, f, adev, step, partial);
    ED_CHECK_LAUNCH();
    return ED_OK;
}

// y = -a[step]*(f/b[step]) + y ; partial sum y^2 -- k_lanczos_update with v = (1/b[step]) f, the
// product the H.v pass forms when it scales f, so the same bits as with a stored v
// 16 B per lane (element pairs), nontemporal: the three streams are read or written once per
// step and do not fit any cache at the sizes this runs on
__global__ __launch_bounds__(kBB) void k_lanczos_update_vf(int64_t n, const double *__restrict__ f,
                                                           double *__restrict__ y, const double *__restrict__ adev,
                                                           const double *__restrict__ bdev, int step,
                                                           double *__restrict__ partial)
{
    typedef double d2 __attribute__((ext_vector_type(2)));
    __shared__ double red[kBB / 64];
    const double c = -adev[step], sc = 1.0 / bdev[step];
    const int64_t base = (int64_t)blockIdx.x * kChunk + 2 * (int64_t)threadIdx.x;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kChunk / (2 * kBB); ++k) {
        const int64_t i = base + (int64_t)k * (2 * kBB);
        if (i + 1 < n) {
            const d2 fv = __builtin_nontemporal_load(reinterpret_cast<const d2 *>(f + i));
            const d2 yv = __builtin_nontemporal_load(reinterpret_cast<const d2 *>(y + i));
            d2 r;
            r.x = fma(c, sc * fv.x, yv.x);
            r.y = fma(c, sc * fv.y, yv.y);
            __builtin_nontemporal_store(r, reinterpret_cast<d2 *>(y + i));
            s = fma(r.x, r.x, s);
            s = fma(r.y, r.y, s);
        } else if (i < n) {
            const double r = fma(c, sc * f[i], y[i]);
            y[i] = r;
            s = fma(r, r, s);
        }
    }
    s = block_sum<kBB>(s, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

int lanczos_update_vfree(int64_t nreal, const double *f, double *y, const double *adev, const double *bdev, int step,
                         double *partial, int64_t *nblocks, hipStream_t st)
{
    const int64_t nb = blas_num_blocks(nreal);
    if (nblocks) *nblocks = nb;
    hipLaunchKernelGGL(k_lanczos_update_vf, dim3((unsigned)nb), dim3(kBB), 0, st, nreal, f, y, adev, bdev, step,
                       partial);
    ED_CHECK_LAUNCH();
    return ED_OK;
}

// f = -a[step]*v + f ; evec = z*v + evec
__global__ __launch_bounds__(kBB) void k_ritz_update(int64_t n, const double *__restrict__ v, double *__restrict__ f,
                                                     double *__restrict__ evec, const double *__restrict__ adev,
                                                     double z, int step)
{
    const double c = -adev[step];
    const int64_t base = (int64_t)blockIdx.x * kChunk + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kChunk / kBB; ++k) {
        const int64_t i = base + (int64_t)k * kBB;
        if (i < n) {
            const double vi = v[i];
            f[i] = fma(c, vi, f[i]);
            evec[i] = fma(z, vi, evec[i]);
        }
    }
}

int ritz_update(int64_t nreal, const double *v, double *f, double *evec, const double *adev, double z, int step,
                hipStream_t st)
{
    const int64_t nb = blas_num_blocks(nreal);
    hipLaunchKernelGGL(k_ritz_update, dim3((unsigned)nb), dim3(kBB), 0, st, nreal, v, f, evec, adev, z, step);
    ED_CHECK_LAUNCH();
    return ED_OK;
}

__global__ __launch_bounds__(kBB) void k_axpy(int64_t n, double alpha, const double *__restrict__ x,
                                              double *__restrict__ y)
{
    const int64_t i = (int64_t)blockIdx.x * kBB + threadIdx.x;
    if (i < n) y[i] = fma(alpha, x[i], y[i]);
}

__global__ __launch_bounds__(kBB) void k_scal(int64_t n, double alpha, double *__restrict__ x)
{
    const int64_t i = (int64_t)blockIdx.x * kBB + threadIdx.x;
    if (i < n) x[i] = alpha * x[i];
}

// x = (1/sqrt(*norm2)) * x
__global__ __launch_bounds__(kBB) void k_scal_invsqrt(int64_t n, const double *__restrict__ norm2,
                                                      double *__restrict__ x)
{
    const double s = 1. / sqrt(*norm2);
    const int64_t i = (int64_t)blockIdx.x * kBB + threadIdx.x;
    if (i < n) x[i] = s * x[i];
}

// minstd_rand (a = 48271, m = 2^31 - 1) with Mersenne reduction
__device__ __forceinline__ uint64_t mulmod31(uint64_t a, uint64_t b)
{
    uint64_t p = a * b;
    uint64_t r = (p & 0x7fffffffULL) + (p >> 31);
    r = (r & 0x7fffffffULL) + (r >> 31);
    return r >= 0x7fffffffULL ? r - 0x7fffffffULL : r;
}

__device__ __forceinline__ uint64_t powmod31(uint64_t base, uint64_t e)
{
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = mulmod31(r, base);
        base = mulmod31(base, base);
        e >>= 1;
    }
    return r;
}

// uniform_real_distribution(-1,1) of one draw: ((x - 1)/(m - 1)) * 2 - 1
__device__ __forceinline__ double lcg_uniform(uint64_t x)
{
    double r = (double)(x - 1);
    r /= 2147483646.0;
    return r * 2.0 + (-1.0);
}

__global__ __launch_bounds__(kBB) void k_random(int64_t n, uint64_t seed, int cplx_, double *__restrict__ v)
{
    const int64_t i = (int64_t)blockIdx.x * kBB + threadIdx.x;
    if (i >= n) return;
    const uint64_t x1 = mulmod31(seed, powmod31(48271ULL, (uint64_t)i + 1));
    if (!cplx_) {
        v[i] = lcg_uniform(x1);
    } else {
        const uint64_t x2 = mulmod31(x1, 48271ULL);
        v[2 * i] = lcg_uniform(x1);
        v[2 * i + 1] = lcg_uniform(x2);
    }
}

} // namespace ed

using namespace ed;

namespace {
struct Scratch {
    double *partial = nullptr;
    double *scal = nullptr;
    int64_t cap = 0;
    // grow-only; released when the thread exits (not during process teardown, ed_thread.hpp)
    ~Scratch()
    {
        if (process_exiting()) return;
        dfree(partial);
        dfree(scal);
    }
    int ensure(int64_t n)
    {
        if (!scal) ED_TRY(dmalloc((void **)&scal, sizeof(double) * 8));
        if (n <= cap) return ED_OK;
        dfree(partial);
        partial = nullptr;
        cap = 0;
        ED_TRY(dmalloc((void **)&partial, sizeof(double) * (size_t)n));
        cap = n;
        return ED_OK;
    }
};
thread_local Scratch g_scratch;
} // namespace

extern "C" int ed_random_vector(int64_t n, uint32_t seed, int complex_vectors, void *d_v, int normalize, void *stream)
{
    if (n < 1 || !d_v) return fail(ED_EINVAL, "ed_random_vector: bad arguments");
    hipStream_t st = stream_of(stream);
    uint64_t s = seed % 2147483647ULL;
    if (s == 0) s = 1; // linear_congruential_engine::seed maps 0 to 1
    hipLaunchKernelGGL(k_random, dim3((unsigned)((n + kBB - 1) / kBB)), dim3(kBB), 0, st, n, s, complex_vectors,
                       (double *)d_v);
    ED_CHECK_LAUNCH();
    if (normalize) return ed_normalize(n, complex_vectors, d_v, nullptr, stream);
    return ED_OK;
}

extern "C" int ed_dot(int64_t n, int complex_vectors, const void *d_a, const void *d_b, double *out, void *stream)
{
    if (n < 1 || !d_a || !d_b || !out) return fail(ED_EINVAL, "ed_dot: bad arguments");
    hipStream_t st = stream_of(stream);
    const int64_t nreal = complex_vectors ? 2 * n : n;
    ED_TRY(g_scratch.ensure(blas_num_blocks(nreal)));
    int64_t nb;
    ED_TRY(dot_partials(nreal, (const double *)d_a, (const double *)d_b, g_scratch.partial, &nb, st));
    ED_TRY(finalize_sum(g_scratch.partial, nb, g_scratch.scal, 0, st));
    out[1] = 0.0;
    if (complex_vectors) {
        const int64_t nb2 = blas_num_blocks(n);
        hipLaunchKernelGGL(k_dot_im, dim3((unsigned)nb2), dim3(kBB), 0, st, n, (const double *)d_a,
                           (const double *)d_b, g_scratch.partial);
        ED_CHECK_LAUNCH();
        ED_TRY(finalize_sum(g_scratch.partial, nb2, g_scratch.scal + 1, 0, st));
    }
    double h[2] = {0, 0};
    ED_HIP(hipMemcpyAsync(h, g_scratch.scal, sizeof(double) * 2, hipMemcpyDeviceToHost, st));
    ED_HIP(hipStreamSynchronize(st));
    out[0] = h[0];
    if (complex_vectors) out[1] = h[1];
    return ED_OK;
}

extern "C" int ed_axpy(int64_t n, int complex_vectors, double alpha, const void *d_x, void *d_y, void *stream)
{
    if (n < 1 || !d_x || !d_y) return fail(ED_EINVAL, "ed_axpy: bad arguments");
    const int64_t nreal = complex_vectors ? 2 * n : n;
    hipLaunchKernelGGL(k_axpy, dim3((unsigned)((nreal + kBB - 1) / kBB)), dim3(kBB), 0, stream_of(stream), nreal,
                       alpha, (const double *)d_x, (double *)d_y);
    ED_CHECK_LAUNCH();
    return ED_OK;
}

extern "C" int ed_scal(int64_t n, int complex_vectors, double alpha, void *d_x, void *stream)
{
    if (n < 1 || !d_x) return fail(ED_EINVAL, "ed_scal: bad arguments");
    const int64_t nreal = complex_vectors ? 2 * n : n;
    hipLaunchKernelGGL(k_scal, dim3((unsigned)((nreal + kBB - 1) / kBB)), dim3(kBB), 0, stream_of(stream), nreal,
                       alpha, (double *)d_x);
    ED_CHECK_LAUNCH();
    return ED_OK;
}

extern "C" int ed_normalize(int64_t n, int complex_vectors, void *d_x, double *norm2_out, void *stream)
{
    if (n < 1 || !d_x) return fail(ED_EINVAL, "ed_normalize: bad arguments");
    hipStream_t st = stream_of(stream);
    const int64_t nreal = complex_vectors ? 2 * n : n;
    ED_TRY(g_scratch.ensure(blas_num_blocks(nreal)));
    int64_t nb;
    ED_TRY(dot_partials(nreal, (const double *)d_x, (const double *)d_x, g_scratch.partial, &nb, st));
    ED_TRY(finalize_sum(g_scratch.partial, nb, g_scratch.scal, 0, st));
    hipLaunchKernelGGL(k_scal_invsqrt, dim3((unsigned)((nreal + kBB - 1) / kBB)), dim3(kBB), 0, st, nreal,
                       g_scratch.scal, (double *)d_x);
    ED_CHECK_LAUNCH();
    if (norm2_out) {
        ED_HIP(hipMemcpyAsync(norm2_out, g_scratch.scal, sizeof(double), hipMemcpyDeviceToHost, st));
        ED_HIP(hipStreamSynchronize(st));
    }
    return ED_OK;
}

// ---- allocation-free dot product ---------------------------------------------------------------
// ed_dot_noalloc takes its block partials from this file-scope device array instead of a workspace
// it allocates: 2^21 partials (16 MiB on every GPU the library runs on, held until the process
// ends), i.e. vectors of at most 2^21 * 2048 = 2^32 real elements.  In host code `scratch` names a
// host-side stand-in; kernels get the device address from hipGetSymbolAddress, looked up once per
// device with that device current.
constexpr int64_t kNoAllocPartials = (int64_t)1 << 21;
static __device__ __align__(256) double scratch[kNoAllocPartials];

namespace {
constexpr int kMaxDevices = 64;
double *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

// the calling thread's current device's copy of `scratch`
int scratch_ptr(double **p)
{
    int dev = 0;
    ED_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(ED_EUNSUPPORTED, "ed_dot_noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *q = nullptr;
        ED_HIP(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = (double *)q;
    }
    *p = g_scratch_ptr[dev];
    return ED_OK;
}

__global__ void k_store_zero(double *p) { *p = 0.0; }
} // namespace

extern "C" int ed_dot_noalloc(int64_t n, int complex_vectors, const void *d_a, const void *d_b, double *d_out,
                              void *stream)
{
    if (n < 1 || !d_a || !d_b || !d_out) return fail(ED_EINVAL, "ed_dot_noalloc: bad arguments");
    const int64_t nreal = complex_vectors ? 2 * n : n;
    if (blas_num_blocks(nreal) > kNoAllocPartials)
        return fail(ED_EINVAL, "ed_dot_noalloc: %lld real elements exceed the fixed partial buffer (at most %lld)",
                    (long long)nreal, (long long)(kNoAllocPartials * kChunk));
    double *part = nullptr;
    ED_TRY(scratch_ptr(&part));
    hipStream_t st = stream_of(stream);
    int64_t nb;
    ED_TRY(dot_partials(nreal, (const double *)d_a, (const double *)d_b, part, &nb, st));
    ED_TRY(finalize_sum(part, nb, d_out, 0, st));
    if (complex_vectors) {
        // stream order: the finalize above has read the partials before k_dot_im overwrites them
        const int64_t nb2 = blas_num_blocks(n);
        hipLaunchKernelGGL(k_dot_im, dim3((unsigned)nb2), dim3(kBB), 0, st, n, (const double *)d_a,
                           (const double *)d_b, part);
        ED_CHECK_LAUNCH();
        ED_TRY(finalize_sum(part, nb2, d_out + 1, 0, st));
    } else {
        hipLaunchKernelGGL(k_store_zero, dim3(1), dim3(1), 0, st, d_out + 1);
        ED_CHECK_LAUNCH();
    }
    return ED_OK;
}

// ed_symeig.hip -- all eigenvalues of a dense real symmetric matrix by a two-stage reduction,
// for the large particle-number blocks of rho_A (ed_entangle.hip).
//
// The reference takes the block spectra from LAPACK/cuSOLVER SVDs (entangle_entropy.h:231-353);
// the one-stage tridiagonal reduction behind such a call (and behind rocSOLVER's dsyevd) reads the
// whole trailing matrix once per column in matrix-vector products, n^3/3 doubles per block, which
// bounds it by HBM bandwidth and kernel launches.  Here:
//   1. dense -> band of half-bandwidth B: blocked Householder QR of each B-column panel below the
//      band (rocSOLVER geqrf + larft), then the two-sided update of the trailing matrix as
//      rank-2B BLAS-3 calls on the lower triangle (X = A V T, W = X - V (T^T V^T X) / 2,
//      A -= V W^T + W V^T: rocBLAS symm, gemm, syr2k);
//   2. band -> tridiagonal by bulge chasing (Lang's algorithm: sweep s annihilates column s below
//      the subdiagonal with one reflector, applied to the next diagonal block; the fill it leaves
//      in the block below is chased down one block at a time, each step annihilating only the
//      first column of the new bulge).  Sweep s's step k runs at time step k + 2 s: the steps of
//      one time step touch disjoint parts of the band, so each time step is one launch with one
//      workgroup per step;
//   3. every eigenvalue of the tridiagonal by Sturm bisection, one thread per eigenvalue index.
// All three stages are orthogonal similarity transformations followed by bisection, so the
// eigenvalues are accurate to a few eps |A| in absolute terms, as those of dsyevd are.
#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "ed_common.hpp"
#include "ed_thread.hpp"

namespace ed {

constexpr int kSyB = 64;   // half-bandwidth of the band form (stage 1 panel width)
constexpr int kSyT = 512;  // threads of a bulge-chasing step

// count symmetric matrices of order n (ld lda, lda * n doubles apart) whose eigenvalues go to w
// (n per matrix, ascending)
struct SymGroup {
    int64_t n, lda;
    int count;
    double *A, *w;
};

// explicit reflector block V (m x kb, ld m: unit diagonal, zero above) from a geqrf panel, whose
// reflector part is then cleared: the panel keeps R, the band part of its columns.  blockIdx.y:
// the matrix of a group (panels sP, blocks sV doubles apart).
__global__ __launch_bounds__(256) void k_sy_vexplicit(int64_t m, int kb, double *__restrict__ P, int64_t lda,
                                                      int64_t sP, double *__restrict__ V, int64_t sV)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= m * kb) return;
    P += (int64_t)blockIdx.y * sP;
    V += (int64_t)blockIdx.y * sV;
    const int64_t r = i % m, c = i / m;
    if (r > c) {
        V[i] = P[r + c * lda];
        P[r + c * lda] = 0.0;
    } else {
        V[i] = r == c ? 1.0 : 0.0;
    }
}

// T of the compact WY form I - V T V^T = H_0 H_1 ... H_{kb-1} (forward, columnwise), from
// G = V^T V: T(j, j) = tau_j, T(0:j, j) = -tau_j T(0:j, 0:j) G(0:j, j).  One workgroup; T is
// kB x kB (ld kB), zero below the diagonal.
template <int KB>
__global__ __launch_bounds__(256) void k_sy_larft(int kb, const double *__restrict__ G, const double *__restrict__ tau,
                                                  double *__restrict__ T)
{
    __shared__ double sT[KB][KB + 1], sg[KB];
    const int t = threadIdx.x;
    G += (int64_t)blockIdx.x * KB * KB; // one workgroup per matrix of the group
    tau += (int64_t)blockIdx.x * KB;
    T += (int64_t)blockIdx.x * KB * KB;
    for (int e = t; e < KB * KB; e += 256) sT[e % KB][e / KB] = 0.0;
    __syncthreads();
    for (int j = 0; j < kb; ++j) {
        if (t < j) sg[t] = G[t + j * KB];
        __syncthreads();
        if (t < j) {
            double acc = 0.0;
            for (int q = t; q < j; ++q) acc = fma(sT[t][q], sg[q], acc); // T upper triangular
            sT[t][j] = -tau[j] * acc;
        }
        if (t == 0) sT[j][j] = tau[j];
        __syncthreads();
    }
    for (int e = t; e < KB * KB; e += 256) T[e] = sT[e % KB][e / KB];
}

// [V | W] and [W | V] side by side (m x 2kb each, ld m) for the rank-2kb update as one GEMM
__global__ __launch_bounds__(256) void k_sy_pair(int64_t m, int kb, const double *__restrict__ V,
                                                 const double *__restrict__ W, int64_t sV, double *__restrict__ VW,
                                                 double *__restrict__ WV, int64_t sVW)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= m * kb) return;
    V += (int64_t)blockIdx.y * sV;
    W += (int64_t)blockIdx.y * sV;
    VW += (int64_t)blockIdx.y * sVW;
    WV += (int64_t)blockIdx.y * sVW;
    const double v = V[i], w = W[i];
    VW[i] = v;
    VW[i + m * kb] = w;
    WV[i] = w;
    WV[i + m * kb] = v;
}

// the other triangle <- one triangle (full storage for the GEMM-based updates)
__global__ __launch_bounds__(256) void k_sy_fill(int64_t n, double *__restrict__ A, int64_t lda, int64_t sA,
                                                 int from_upper)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n * n) return;
    A += (int64_t)blockIdx.y * sA;
    const int64_t r = i % n, c = i / n;
    if (from_upper ? r > c : r < c) A[r + c * lda] = A[c + r * lda];
}

// Householder reflector of x[0..L-1] in LDS (one wavefront): x <- v with v[0] = 1, returns tau and
// beta, (I - tau v v^T) x = beta e_0 (LAPACK's dlarfg convention, beta = -sign(x0) |x|)
__device__ __forceinline__ void sy_house(double *x, int L, double *tau, double *beta)
{
    const int lane = threadIdx.x & 63;
    double s = 0.0;
    for (int i = 1 + lane; i < L; i += 64) s = fma(x[i], x[i], s);
    s = wave_sum(s);
    const double alpha = x[0];
    if (s == 0.0) {
        *tau = 0.0;
        *beta = alpha;
        if (lane == 0) x[0] = 1.0;
        return;
    }
    const double nrm = std::sqrt(alpha * alpha + s);
    const double b = alpha >= 0.0 ? -nrm : nrm;
    const double sc = 1.0 / (alpha - b);
    for (int i = 1 + lane; i < L; i += 64) x[i] *= sc;
    if (lane == 0) x[0] = 1.0;
    *tau = (b - alpha) / b;
    *beta = b;
}

// D <- H D H for the L x L symmetric block in LDS (full, ld kSyB + 1), H = I - tau v v^T
__device__ __forceinline__ void sy_two_sided(double (*D)[kSyB + 1], int L, const double *v, double tau, double *p,
                                             double *red)
{
    const int t = threadIdx.x;
    constexpr int PARTS = kSyT / kSyB; // 4 partial sums per row
    // p = tau D v
    {
        const int i = t % kSyB, part = t / kSyB;
        double acc = 0.0;
        if (i < L)
            for (int j = part; j < L; j += PARTS) acc = fma(D[i][j], v[j], acc);
        red[part * kSyB + i] = acc;
    }
    __syncthreads();
    if (t < kSyB) {
        double acc = 0.0;
        if (t < L)
            for (int q = 0; q < PARTS; ++q) acc += red[q * kSyB + t];
        p[t] = tau * acc;
    }
    __syncthreads();
    // w = p - (tau / 2) (p . v) v
    if (t < 64) {
        double s = 0.0;
        for (int i = t; i < L; i += 64) s = fma(p[i], v[i], s);
        s = wave_sum(s);
        const double c = 0.5 * tau * s;
        for (int i = t; i < L; i += 64) p[i] = p[i] - c * v[i];
    }
    __syncthreads();
    for (int e = t; e < L * L; e += kSyT) {
        const int i = e % L, j = e / L;
        D[i][j] -= v[i] * p[j] + p[i] * v[j];
    }
    __syncthreads();
}

// A matrix of the bulge chasing: order n, lower triangle of A (ld lda, half-bandwidth kSyB), refl
// per sweep its last reflector (kSyB entries of v, then tau); its steps take the workgroups
// slot0 .. slot0 + slots - 1 of every launch; kmax0 = steps after the first of sweep 0.
struct SbMat {
    double *A;
    int64_t n, lda;
    double *refl;
    int64_t slot0, slots, kmax0;
};

// One time step `ts` of the bulge chasing over several matrices: the workgroups of matrix q run
// step k = ts - 2 s of the sweeps s from max(0, ceil((ts - kmax0) / 2)) on.
__global__ __launch_bounds__(kSyT) void k_sb2st_step(const SbMat *__restrict__ mats, int nmats, int64_t ts)
{
    constexpr int B = kSyB;
    __shared__ double D[B][B + 1];
    __shared__ double v[B], v2[B], p[B], red[kSyT];
    __shared__ double sc[4];
    int lo = 0, hi = nmats - 1; // the last matrix with slot0 <= blockIdx.x
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (mats[mid].slot0 <= (int64_t)blockIdx.x) lo = mid;
        else hi = mid - 1;
    }
    const SbMat &M = mats[lo];
    const int64_t n = M.n, lda = M.lda;
    double *__restrict__ A = M.A;
    const int64_t q = (int64_t)blockIdx.x - M.slot0;
    if (q >= M.slots) return;
    const int64_t s = std::max<int64_t>(0, (ts - M.kmax0 + 1) / 2) + q, k = ts - 2 * s;
    if (s > n - 3 || k < 0) return;
    if (k > 0 && s + k * B > n - 2) return; // the sweep has ended
    const int t = threadIdx.x;
    double *refl = M.refl;
    double *rv = refl + s * (B + 1);
    if (k == 0) {
        // annihilate column s below the subdiagonal
        const int64_t st = s + 1, ed = std::min<int64_t>(s + B, n - 1);
        const int L = (int)(ed - st + 1);
        if (t < L) v[t] = A[(st + t) + s * lda];
        __syncthreads();
        if (t < 64) {
            double tau, beta;
            sy_house(v, L, &tau, &beta);
            if (t == 0) {
                sc[0] = tau;
                sc[1] = beta;
            }
        }
        __syncthreads();
        const double tau = sc[0];
        if (t < L) A[(st + t) + s * lda] = t == 0 ? sc[1] : 0.0;
        // the lower triangle column by column (contiguous in memory), mirrored in LDS
        for (int e = t; e < L * L; e += kSyT) {
            const int i = e % L, j = e / L;
            if (i >= j) D[i][j] = A[(st + i) + (st + j) * lda];
        }
        __syncthreads();
        for (int e = t; e < L * L; e += kSyT) {
            const int i = e % L, j = e / L;
            if (i < j) D[i][j] = D[j][i];
        }
        __syncthreads();
        sy_two_sided(D, L, v, tau, p, red);
        for (int e = t; e < L * L; e += kSyT) {
            const int i = e % L, j = e / L;
            if (i >= j) A[(st + i) + (st + j) * lda] = D[i][j];
        }
        if (t < B) rv[t] = t < L ? v[t] : 0.0;
        if (t == 0) rv[B] = tau;
        return;
    }
    // chase: the previous reflector (columns st..ed) applied from the right to the block below,
    // a new one from that block's first column, applied from the left and to the next diagonal block.
    // The block below, C, stays in registers: thread (i = t % B, part = t / B) holds row i's
    // columns j = part + PARTS q, so that only the diagonal block takes LDS (four workgroups per CU).
    constexpr int PARTS = kSyT / B, NQ = B / PARTS;
    const int64_t st = s + 1 + (k - 1) * B, ed = std::min<int64_t>(s + k * B, n - 1);
    const int64_t j1 = s + k * B + 1, j2 = std::min<int64_t>(s + (k + 1) * B, n - 1);
    const int Lp = (int)(ed - st + 1), Lc = (int)(j2 - j1 + 1);
    const int ri = t % B, part = t / B;
    // every load of the step up front: the reflector, the block below and the next diagonal block
    // (its lower triangle, contiguous columns, mirrored in LDS below)
    if (t < B) v[t] = t < Lp ? rv[t] : 0.0; // zero past Lp: the products below run over all B columns
    if (t == 0) sc[0] = rv[B];
    double c[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int j = part + PARTS * q;
        c[q] = (ri < Lc && j < Lp) ? A[(j1 + ri) + (st + j) * lda] : 0.0;
    }
    for (int e = t; e < Lc * Lc; e += kSyT) {
        const int i = e % Lc, j = e / Lc;
        if (i >= j) D[i][j] = A[(j1 + i) + (j1 + j) * lda];
    }
    __syncthreads();
    for (int e = t; e < Lc * Lc; e += kSyT) {
        const int i = e % Lc, j = e / Lc;
        if (i < j) D[i][j] = D[j][i];
    }
    const double tau = sc[0];
    // u = C v; C -= tau u v^T
    {
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) acc = fma(c[q], v[part + PARTS * q], acc); // c is 0 past Lp
        red[part * B + ri] = acc;
    }
    __syncthreads();
    if (t < B) {
        double acc = 0.0;
        if (t < Lc)
            for (int q = 0; q < PARTS; ++q) acc += red[q * B + t];
        p[t] = tau * acc;
    }
    __syncthreads();
    {
        const double pi = p[ri];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int j = part + PARTS * q;
            if (ri < Lc && j < Lp) c[q] -= pi * v[j];
        }
    }
    // new reflector from C[:, 0] (held by the threads of part 0, q = 0)
    if (part == 0 && ri < Lc) v2[ri] = c[0];
    __syncthreads();
    if (t < 64) {
        double tau2, beta2;
        sy_house(v2, Lc, &tau2, &beta2);
        if (t == 0) {
            sc[2] = tau2;
            sc[3] = beta2;
        }
    }
    __syncthreads();
    const double tau2 = sc[2];
    // C[:, 1:] <- H2 C[:, 1:]: r_j = v2 . C[:, j] summed over the wavefront holding column j;
    // C[i][j] -= tau2 v2_i r_j
    {
        const double vi = ri < Lc ? v2[ri] : 0.0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const double r = wave_sum(vi * c[q]);
            const int j = part + PARTS * q;
            if (ri == 0) p[j] = tau2 * r;
        }
    }
    __syncthreads();
    {
        const double vi = ri < Lc ? v2[ri] : 0.0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int j = part + PARTS * q;
            if (ri < Lc && j < Lp) {
                const double cv = j == 0 ? (ri == 0 ? sc[3] : 0.0) : c[q] - vi * p[j];
                A[(j1 + ri) + (st + j) * lda] = cv;
            }
        }
    }
    // the next diagonal block, two-sided (p is free again once every thread has read it)
    __syncthreads();
    sy_two_sided(D, Lc, v2, tau2, p, red);
    for (int e = t; e < Lc * Lc; e += kSyT) {
        const int i = e % Lc, j = e / Lc;
        if (i >= j) A[(j1 + i) + (j1 + j) * lda] = D[i][j];
    }
    if (t < B) rv[t] = t < Lc ? v2[t] : 0.0;
    if (t == 0) rv[B] = tau2;
}

// diagonal and subdiagonal of the tridiagonal left in A
__global__ __launch_bounds__(256) void k_sy_tri_extract(int64_t n, const double *__restrict__ A, int64_t lda,
                                                        double *__restrict__ d, double *__restrict__ e)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    d[i] = A[i + i * lda];
    e[i] = i + 1 < n ? A[(i + 1) + i * lda] : 0.0;
}

// Sturm count of T - x: number of eigenvalues below x (LDL^T pivots; a zero pivot is replaced by
// the smallest normal double)
__device__ __forceinline__ int64_t sy_sturm(int64_t n, const double *__restrict__ d, const double *__restrict__ e,
                                            double x)
{
    int64_t c = 0;
    double q = d[0] - x;
    if (q < 0.0) ++c;
    for (int64_t i = 1; i < n; ++i) {
        if (q == 0.0) q = std::numeric_limits<double>::min();
        q = (d[i] - x) - e[i - 1] * e[i - 1] / q;
        if (q < 0.0) ++c;
    }
    return c;
}

// a tridiagonal of the bisection stage: d, e (n each), eigenvalues into w; its eigenvalue
// indices are the threads off .. off + n - 1 of the launch; [lo, hi] a Gershgorin interval
struct TriMat {
    const double *d, *e;
    double *w;
    int64_t n, off;
    double lo, hi;
};

// w[i] = the i-th smallest eigenvalue, by bisection of [lo, hi] until the bracket is within 4 eps
// of its value or eps * max(|lo|, |hi|) / 2^20 wide
__global__ __launch_bounds__(256) void k_sy_tri_bisect(const TriMat *__restrict__ mats, int nmats, int64_t total)
{
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= total) return;
    int a = 0, b = nmats - 1; // the last matrix with off <= g
    while (a < b) {
        const int mid = (a + b + 1) / 2;
        if (mats[mid].off <= g) a = mid;
        else b = mid - 1;
    }
    const TriMat &M = mats[a];
    const int64_t i = g - M.off;
    const double eps = std::numeric_limits<double>::epsilon();
    const double floor_w = eps * std::fmax(std::fabs(M.lo), std::fabs(M.hi)) * 0x1p-20;
    double lo = M.lo, hi = M.hi;
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (!(mid > lo && mid < hi)) break;
        if (hi - lo <= std::fmax(4.0 * eps * std::fmax(std::fabs(lo), std::fabs(hi)), floor_w)) break;
        if (sy_sturm(M.n, M.d, M.e, mid) > i) hi = mid;
        else lo = mid;
    }
    M.w[i] = 0.5 * (lo + hi);
}

// stage 1 for the count matrices of one group (same order n, ld lda, stride sA), strided batched
static int band_reduce_group(rocblas_handle h, hipStream_t st, int64_t n, int64_t lda, int64_t sA, int count,
                             double *A, double *V, double *X, double *W, double *VW, double *WV, double *T,
                             double *Y, double *Z, double *tau)
{
    constexpr int B = kSyB;
    const rocblas_int ild = (rocblas_int)lda, cnt = count;
    const int64_t sV = n * B, sVW = 2 * n * B, sBB = B * B;
    for (int64_t k = 0; k + B < n; k += B) {
        const int64_t m = n - k - B;
        const int kb = (int)std::min<int64_t>(B, m);
        double *P = A + (k + B) + k * lda; // m x B panels below the band
        double *A22 = A + (k + B) + (k + B) * lda;
        const rocblas_int im = (rocblas_int)m;
        const double one = 1.0, zero = 0.0, mhalf = -0.5, mone = -1.0;
        rocblas_status rs = rocsolver_dgeqrf_strided_batched(h, im, B, P, ild, sA, tau, B, cnt);
        if (rs != rocblas_status_success) return fail(ED_EHIP, "two-stage eigenvalues: panel QR failed (%d)", (int)rs);
        const dim3 gmk((unsigned)((m * kb + 255) / 256), (unsigned)count);
        hipLaunchKernelGGL(k_sy_vexplicit, gmk, dim3(256), 0, st, m, kb, P, lda, sA, V, sV);
        ED_CHECK_LAUNCH();
        // T from G = V^T V
        rs = rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, kb, kb, im, &one, V,
                                           im, sV, V, im, sV, &zero, Y, B, sBB, cnt);
        if (rs != rocblas_status_success) return fail(ED_EHIP, "two-stage eigenvalues: V^T V failed (%d)", (int)rs);
        hipLaunchKernelGGL(k_sy_larft<B>, dim3((unsigned)count), dim3(256), 0, st, kb, (const double *)Y,
                           (const double *)tau, T);
        ED_CHECK_LAUNCH();
        // X = A22 V T
        rs = rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, im, kb, im, &one, A22,
                                           ild, sA, V, im, sV, &zero, W, im, sV, cnt);
        if (rs == rocblas_status_success)
            rs = rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, im, kb, kb, &one, W,
                                               im, sV, T, B, sBB, &zero, X, im, sV, cnt);
        // Y = V^T X, Z = T^T Y, W = X - V Z / 2
        if (rs == rocblas_status_success)
            rs = rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, kb, kb, im,
                                               &one, V, im, sV, X, im, sV, &zero, Y, B, sBB, cnt);
        if (rs == rocblas_status_success)
            rs = rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, kb, kb, kb,
                                               &one, T, B, sBB, Y, B, sBB, &zero, Z, B, sBB, cnt);
        if (rs == rocblas_status_success)
            ED_HIP(hipMemcpyAsync(W, X, sizeof(double) * (size_t)sV * count, hipMemcpyDeviceToDevice, st));
        if (rs == rocblas_status_success)
            rs = rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, im, kb, kb, &mhalf, V,
                                               im, sV, Z, B, sBB, &one, W, im, sV, cnt);
        // A22 -= [V W] [W V]^T = V W^T + W V^T (both triangles)
        if (rs == rocblas_status_success) {
            hipLaunchKernelGGL(k_sy_pair, gmk, dim3(256), 0, st, m, kb, (const double *)V, (const double *)W, sV, VW, WV,
                               sVW);
            ED_CHECK_LAUNCH();
            rs = rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_transpose, im, im, 2 * kb,
                                               &mone, VW, im, sVW, WV, im, sVW, &one, A22, ild, sA, cnt);
        }
        if (rs != rocblas_status_success)
            return fail(ED_EHIP, "two-stage eigenvalues: trailing update failed (%d)", (int)rs);
    }
    return ED_OK;
}

int sym_eigvals_groups(rocblas_handle h, hipStream_t st, const std::vector<SymGroup> &groups, bool upper_only)
{
    constexpr int B = kSyB;
    int64_t maxwork = 1, total = 0, nmat = 0;
    for (const SymGroup &g : groups) {
        if (g.n < 1 || g.count < 1 || !g.A || !g.w || g.lda < g.n || g.n > 0x7fffffffLL)
            return fail(ED_EINVAL, "two-stage eigenvalues: bad group");
        maxwork = std::max(maxwork, g.n * (int64_t)g.count);
        total += g.n * g.count;
        nmat += g.count;
    }
    if (nmat == 0) return ED_OK;
    rocblas_set_stream(h, st);
    rocblas_pointer_mode pm;
    rocblas_get_pointer_mode(h, &pm);
    rocblas_set_pointer_mode(h, rocblas_pointer_mode_host);
    struct PmGuard {
        rocblas_handle h;
        rocblas_pointer_mode pm;
        ~PmGuard() { rocblas_set_pointer_mode(h, pm); }
    } pm_guard{h, pm};
    int64_t maxcount = 1;
    for (const SymGroup &g : groups) maxcount = std::max<int64_t>(maxcount, g.count);
    static thread_local Arena t_sy_arena;
    Scratch S(t_sy_arena, st);
    const size_t bnb = Scratch::round(sizeof(double) * (size_t)maxwork * B),
                 bbb = Scratch::round(sizeof(double) * B * B * (size_t)maxcount),
                 bb = Scratch::round(sizeof(double) * B * (size_t)maxcount),
                 bt = Scratch::round(sizeof(double) * (size_t)total),
                 br = Scratch::round(sizeof(double) * (size_t)total * (B + 1)),
                 bsb = Scratch::round(sizeof(SbMat) * (size_t)nmat), btm = Scratch::round(sizeof(TriMat) * (size_t)nmat);
    ED_TRY(S.reserve(7 * bnb + 3 * bbb + bb + 3 * bt + br + bsb + btm));
    double *V = (double *)S.take(bnb), *X = (double *)S.take(bnb), *W = (double *)S.take(bnb);
    double *VW = (double *)S.take(2 * bnb), *WV = (double *)S.take(2 * bnb);
    double *T = (double *)S.take(bbb), *Y = (double *)S.take(bbb), *Z = (double *)S.take(bbb);
    double *tau = (double *)S.take(bb);
    double *dd = (double *)S.take(bt), *de = (double *)S.take(bt), *dw = (double *)S.take(bt);
    double *refl = (double *)S.take(br);
    SbMat *dsb = (SbMat *)S.take(bsb);
    TriMat *dtm = (TriMat *)S.take(btm);
    (void)dw;
    // 1. dense -> band, group by group (full storage: the trailing updates are plain GEMMs)
    for (const SymGroup &g : groups) {
        const int64_t sA = g.lda * g.n;
        hipLaunchKernelGGL(k_sy_fill, dim3((unsigned)((g.n * g.n + 255) / 256), (unsigned)g.count), dim3(256), 0, st,
                           g.n, g.A, g.lda, sA, upper_only ? 1 : 0);
        ED_CHECK_LAUNCH();
        ED_TRY(band_reduce_group(h, st, g.n, g.lda, sA, g.count, g.A, V, X, W, VW, WV, T, Y, Z, tau));
    }
    // 2. band -> tridiagonal, every matrix in each launch (larger matrices first: the grid of a
    // time step is the prefix of matrices still chasing)
    std::vector<SbMat> sb;
    std::vector<TriMat> tm;
    {
        int64_t roff = 0, off = 0;
        for (const SymGroup &g : groups)
            for (int c = 0; c < g.count; ++c) {
                SbMat M;
                M.A = g.A + (int64_t)c * g.lda * g.n;
                M.n = g.n;
                M.lda = g.lda;
                M.refl = refl + roff;
                M.kmax0 = g.n >= 2 ? (g.n - 2) / B : 0;
                M.slots = M.kmax0 / 2 + 2;
                M.slot0 = 0;
                sb.push_back(M);
                TriMat Tm;
                Tm.d = dd + off;
                Tm.e = de + off;
                Tm.w = g.w + (int64_t)c * g.n;
                Tm.n = g.n;
                Tm.off = off;
                Tm.lo = Tm.hi = 0.0;
                tm.push_back(Tm);
                roff += g.n * (B + 1);
                off += g.n;
            }
    }
    std::vector<int> ord(sb.size());
    for (size_t i = 0; i < ord.size(); ++i) ord[i] = (int)i;
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return sb[x].n > sb[y].n; });
    std::vector<SbMat> sbo;
    std::vector<int64_t> tsmax, prefix;
    int64_t slot = 0;
    for (int i : ord) {
        SbMat M = sb[i];
        if (M.n < 3) continue; // already tridiagonal
        M.slot0 = slot;
        slot += M.slots;
        sbo.push_back(M);
        tsmax.push_back(2 * (M.n - 3) + M.kmax0);
        prefix.push_back(slot);
    }
    if (!sbo.empty()) {
        ED_HIP(hipMemcpyAsync(dsb, sbo.data(), sizeof(SbMat) * sbo.size(), hipMemcpyHostToDevice, st));
        const int64_t tsend = tsmax[0];
        size_t active = sbo.size();
        for (int64_t ts = 0; ts <= tsend; ++ts) {
            while (active > 1 && tsmax[active - 1] < ts) --active;
            hipLaunchKernelGGL(k_sb2st_step, dim3((unsigned)prefix[active - 1]), dim3(kSyT), 0, st,
                               (const SbMat *)dsb, (int)active, ts);
        }
        ED_CHECK_LAUNCH();
    }
    // 3. eigenvalues of the tridiagonals
    for (size_t q = 0; q < sb.size(); ++q) {
        hipLaunchKernelGGL(k_sy_tri_extract, dim3((unsigned)((sb[q].n + 255) / 256)), dim3(256), 0, st, sb[q].n,
                           (const double *)sb[q].A, sb[q].lda, dd + tm[q].off, de + tm[q].off);
    }
    ED_CHECK_LAUNCH();
    std::vector<double> hd((size_t)total), he((size_t)total);
    ED_HIP(hipMemcpyAsync(hd.data(), dd, sizeof(double) * (size_t)total, hipMemcpyDeviceToHost, st));
    ED_HIP(hipMemcpyAsync(he.data(), de, sizeof(double) * (size_t)total, hipMemcpyDeviceToHost, st));
    ED_HIP(hipStreamSynchronize(st));
    for (TriMat &Tm : tm) {
        const double *d = hd.data() + Tm.off, *e = he.data() + Tm.off;
        double lo = d[0], hi = d[0];
        for (int64_t i = 0; i < Tm.n; ++i) {
            const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i < Tm.n - 1 ? std::fabs(e[i]) : 0.0);
            lo = std::fmin(lo, d[i] - r);
            hi = std::fmax(hi, d[i] + r);
        }
        const double pad = 4.0 * std::numeric_limits<double>::epsilon() * std::fmax(std::fabs(lo), std::fabs(hi)) + 1e-300;
        Tm.lo = lo - pad;
        Tm.hi = hi + pad;
    }
    ED_HIP(hipMemcpyAsync(dtm, tm.data(), sizeof(TriMat) * tm.size(), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_sy_tri_bisect, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, (const TriMat *)dtm,
                       (int)tm.size(), total);
    ED_CHECK_LAUNCH();
    ED_HIP(hipStreamSynchronize(st)); // the host arrays above are read by the uploads
    return ED_OK;
}

int sym_eigvals_two_stage(rocblas_handle h, hipStream_t st, int64_t n, double *A, int64_t lda, bool upper_only,
                          double *w)
{
    if (n < 1 || !A || !w || lda < n) return fail(ED_EINVAL, "two-stage eigenvalues: bad arguments");
    std::vector<SymGroup> g(1);
    g[0] = SymGroup{n, lda, 1, A, w};
    return sym_eigvals_groups(h, st, g, upper_only);
}

} // namespace ed

using namespace ed;

// All eigenvalues (ascending) of a real symmetric n x n matrix (column-major, ld lda, lower triangle
// read and overwritten; upper = 1: the upper triangle instead, copied down first) by the two-stage
// reduction above.  Measurement and test entry point of the routine ed_entangle.hip uses for its
// large blocks; no counterpart in the reference.
extern "C" int ed_sym_eigvals(int64_t n, double *d_A, int64_t lda, int upper, double *d_w, void *stream)
{
    rocblas_handle h = thread_blas_handle();
    if (!h) return fail(ED_EHIP, "ed_sym_eigvals: rocblas handle creation failed");
    const hipStream_t st = stream_of(stream);
    ED_TRY(sym_eigvals_two_stage(h, st, n, d_A, lda, upper != 0, d_w));
    ED_HIP(hipStreamSynchronize(st));
    return ED_OK;
}

// ed_eigs.hip -- the nev lowest eigenpairs of a Hubbard operator (the role of
// cuSpectra::SymEigsSolver<cuSpectra::Hubbard<T>> with SortRule::SmallestAlge,
// MBL_test/include/SymEigsSolver.cuh:36-475 and ArnoldiOp.cuh:132-219, used by the
// TotalSpinExcited drivers), and the host-array H.v of cuSpectra::Hubbard<T>::perform_op
// (HubMatrix_TotalSpin.cuh:181-322).
//
// Method: thick-restart Lanczos (Wu & Simon, SIAM J. Matrix Anal. Appl. 22 (2000) 602) with
// full reorthogonalisation (two classical Gram-Schmidt passes against the whole basis every
// step: "twice is enough"), which computes the
// same invariant subspace as the reference's implicitly restarted Lanczos (ARPACK dsaupd
// logic) but restarts by keeping Ritz vectors instead of applying shifts.  Convergence test of
// the reference: residual |beta_m y_{m,i}| <= tol * max(eps^(2/3), |theta_i|)
// (SymEigsSolver.cuh:124-141).  The basis stays in HBM (ncv+1 vectors); the projected matrix
// (at most ncv x ncv) is diagonalised on the host by cyclic Jacobi.
//
// Kernels: k_proj (block-partial <V_j, w> for all j in one launch, grid = chunks x vectors,
// fixed-order reductions), k_proj_fin, k_subtract (w -= sum_j V_j c_j), k_restart (V <- V Y
// in place, one element per thread with the m old values in registers).
#include <algorithm>
#include <cmath>
#include <vector>

#include "ed_common.hpp"
#include "ed_hubbard.hpp"

namespace ed {

constexpr int kEB = 256;
constexpr int kEChunk = kEB * 8;
constexpr int kMaxNcv = 64;

// partial[j * nb + b] = (Re, Im) of sum over chunk b of conj(V_j) w
template <bool CPLX>
__global__ __launch_bounds__(kEB) void k_proj(int64_t n, int64_t ld, const double *__restrict__ V,
                                              const double *__restrict__ w, double *__restrict__ partial, int64_t nb)
{
    __shared__ double red[kEB / 64];
    const int j = blockIdx.y;
    const double *v = V + (int64_t)j * ld;
    const int64_t base = (int64_t)blockIdx.x * kEChunk + threadIdx.x;
    double sr = 0.0, si = 0.0;
#pragma unroll
    for (int k = 0; k < kEChunk / kEB; ++k) {
        const int64_t i = base + (int64_t)k * kEB;
        if (i < n) {
            if (CPLX) {
                const double vr = v[2 * i], vi = v[2 * i + 1], wr = w[2 * i], wi = w[2 * i + 1];
                sr = fma(vr, wr, fma(vi, wi, sr));
                si = fma(vr, wi, fma(-vi, wr, si));
            } else {
                sr = fma(v[i], w[i], sr);
            }
        }
    }
    sr = block_sum<kEB>(sr, red);
    if (CPLX) si = block_sum<kEB>(si, red);
    if (threadIdx.x == 0) {
        partial[2 * ((int64_t)j * nb + blockIdx.x)] = sr;
        partial[2 * ((int64_t)j * nb + blockIdx.x) + 1] = si;
    }
}

// c[2j], c[2j+1] = sum_b partial[j][b] (one block per j, fixed order)
__global__ __launch_bounds__(256) void k_proj_fin(const double *__restrict__ partial, int64_t nb,
                                                  double *__restrict__ c)
{
    __shared__ double red[4];
    const int j = blockIdx.x;
    double sr = 0.0, si = 0.0;
    for (int64_t b = threadIdx.x; b < nb; b += 256) {
        sr += partial[2 * ((int64_t)j * nb + b)];
        si += partial[2 * ((int64_t)j * nb + b) + 1];
    }
    sr = block_sum<256>(sr, red);
    si = block_sum<256>(si, red);
    if (threadIdx.x == 0) {
        c[2 * j] = sr;
        c[2 * j + 1] = si;
    }
}

// w -= sum_{j<m} V_j c_j
template <bool CPLX>
__global__ __launch_bounds__(kEB) void k_subtract(int64_t n, int64_t ld, int m, const double *__restrict__ V,
                                                  const double *__restrict__ c, double *__restrict__ w)
{
    const int64_t i = (int64_t)blockIdx.x * kEB + threadIdx.x;
    if (i >= n) return;
    if (CPLX) {
        double wr = w[2 * i], wi = w[2 * i + 1];
        for (int j = 0; j < m; ++j) {
            const double vr = V[(int64_t)j * ld + 2 * i], vi = V[(int64_t)j * ld + 2 * i + 1];
            const double cr = c[2 * j], ci = c[2 * j + 1];
            wr -= vr * cr - vi * ci;
            wi -= vr * ci + vi * cr;
        }
        w[2 * i] = wr;
        w[2 * i + 1] = wi;
    } else {
        double wr = w[i];
        for (int j = 0; j < m; ++j) wr = fma(-c[2 * j], V[(int64_t)j * ld + i], wr);
        w[i] = wr;
    }
}

// in place: V_k'[i] = sum_{j<m} Y[j * k + k'] V_j[i] for k' < k (k <= m <= kMaxNcv)
template <bool CPLX>
__global__ __launch_bounds__(kEB) void k_restart(int64_t n, int64_t ld, int m, int k, double *__restrict__ V,
                                                 const double *__restrict__ Y)
{
    const int64_t i = (int64_t)blockIdx.x * kEB + threadIdx.x;
    if (i >= n) return;
    constexpr int NC = CPLX ? 2 : 1;
    double x[kMaxNcv][NC];
#pragma unroll
    for (int j = 0; j < kMaxNcv; ++j) {
        if (j >= m) break; // uniform
#pragma unroll
        for (int q = 0; q < NC; ++q) x[j][q] = V[(int64_t)j * ld + NC * i + q];
    }
    for (int kk = 0; kk < k; ++kk) {
        double o[NC] = {};
#pragma unroll
        for (int j = 0; j < kMaxNcv; ++j) {
            if (j >= m) break;
            const double y = Y[j * k + kk];
#pragma unroll
            for (int q = 0; q < NC; ++q) o[q] = fma(y, x[j][q], o[q]);
        }
#pragma unroll
        for (int q = 0; q < NC; ++q) V[(int64_t)kk * ld + NC * i + q] = o[q];
    }
}

// cyclic Jacobi: A (m x m, row-major, symmetric) -> ascending eigenvalues w, eigenvectors
// in the columns of Z (row-major, Z[r * m + c])
static void jacobi_eig(int m, std::vector<double> A, std::vector<double> &w, std::vector<double> &Z)
{
    Z.assign((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) Z[(size_t)i * m + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int p = 0; p < m; ++p)
            for (int q = 0; q < m; ++q) {
                const double a = A[(size_t)p * m + q] * A[(size_t)p * m + q];
                tot += a;
                if (p != q) off += a;
            }
        if (off <= 1e-32 * tot || off == 0.0) break;
        for (int p = 0; p < m - 1; ++p)
            for (int q = p + 1; q < m; ++q) {
                const double apq = A[(size_t)p * m + q];
                if (apq == 0.0) continue;
                const double app = A[(size_t)p * m + p], aqq = A[(size_t)q * m + q];
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int r = 0; r < m; ++r) { // A <- A J (columns p, q)
                    const double arp = A[(size_t)r * m + p], arq = A[(size_t)r * m + q];
                    A[(size_t)r * m + p] = c * arp - s * arq;
                    A[(size_t)r * m + q] = s * arp + c * arq;
                }
                for (int r = 0; r < m; ++r) { // A <- J^T A (rows p, q)
                    const double apr = A[(size_t)p * m + r], aqr = A[(size_t)q * m + r];
                    A[(size_t)p * m + r] = c * apr - s * aqr;
                    A[(size_t)q * m + r] = s * apr + c * aqr;
                }
                for (int r = 0; r < m; ++r) { // Z <- Z J
                    const double zrp = Z[(size_t)r * m + p], zrq = Z[(size_t)r * m + q];
                    Z[(size_t)r * m + p] = c * zrp - s * zrq;
                    Z[(size_t)r * m + q] = s * zrp + c * zrq;
                }
            }
    }
    std::vector<int> idx(m);
    for (int i = 0; i < m; ++i) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](int a, int b) { return A[(size_t)a * m + a] < A[(size_t)b * m + b]; });
    std::vector<double> Zs((size_t)m * m);
    w.resize(m);
    for (int c = 0; c < m; ++c) {
        w[c] = A[(size_t)idx[c] * m + idx[c]];
        for (int r = 0; r < m; ++r) Zs[(size_t)r * m + c] = Z[(size_t)r * m + idx[c]];
    }
    Z.swap(Zs);
}

struct EigWork {
    double *V = nullptr;   // (ncv + 1) vectors of ld doubles
    double *part = nullptr; // 2 * ncv * nb
    double *c = nullptr;    // 2 * ncv
    double *Y = nullptr;    // ncv * ncv
    double *nrm = nullptr;  // 1
    ~EigWork()
    {
        dfree(V);
        dfree(part);
        dfree(c);
        dfree(Y);
        dfree(nrm);
    }
};

template <bool CPLX>
static int proj_subtract(int64_t n, int64_t ld, int m, const double *V, double *w, EigWork &W, int64_t nb,
                         double *c_host, hipStream_t st)
{
    hipLaunchKernelGGL((k_proj<CPLX>), dim3((unsigned)nb, (unsigned)m), dim3(kEB), 0, st, n, ld, V, (const double *)w,
                       W.part, nb);
    ED_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_proj_fin, dim3((unsigned)m), dim3(256), 0, st, (const double *)W.part, nb, W.c);
    ED_CHECK_LAUNCH();
    hipLaunchKernelGGL((k_subtract<CPLX>), dim3((unsigned)((n + kEB - 1) / kEB)), dim3(kEB), 0, st, n, ld, m, V,
                       (const double *)W.c, w);
    ED_CHECK_LAUNCH();
    ED_HIP(hipMemcpyAsync(c_host, W.c, sizeof(double) * 2 * m, hipMemcpyDeviceToHost, st));
    return ED_OK;
}

static int norm2(int64_t nreal, const double *x, EigWork &W, int64_t *nbv, double *out, hipStream_t st)
{
    ED_TRY(dot_partials(nreal, x, x, W.part, nbv, st));
    ED_TRY(finalize_sum(W.part, *nbv, W.nrm, 1, st));
    ED_HIP(hipMemcpyAsync(out, W.nrm, sizeof(double), hipMemcpyDeviceToHost, st));
    ED_HIP(hipStreamSynchronize(st));
    return ED_OK;
}

template <bool CPLX>
static int eigs_run(const Hubbard *h, int nev, int ncv, int maxit, double tol, uint32_t seed, double *evals,
                    void *d_evecs, int *nconv_out, int *niter_out, hipStream_t st)
{
    const int64_t n = h->args.dim_up * h->args.dim_dn;
    const int64_t nreal = CPLX ? 2 * n : n;
    const int64_t ld = nreal;
    const int m = (int)std::min<int64_t>(ncv, n);
    const int64_t nb = (n + kEChunk - 1) / kEChunk;
    const int64_t nbblas = blas_num_blocks(nreal);
    EigWork W;
    ED_TRY(dmalloc((void **)&W.V, sizeof(double) * (size_t)ld * (m + 1)));
    ED_TRY(dmalloc((void **)&W.part, sizeof(double) * 2 * (size_t)std::max<int64_t>(m * nb, nbblas)));
    ED_TRY(dmalloc((void **)&W.c, sizeof(double) * 2 * m));
    ED_TRY(dmalloc((void **)&W.Y, sizeof(double) * m * m));
    ED_TRY(dmalloc((void **)&W.nrm, sizeof(double)));
    auto vec = [&](int j) { return W.V + (int64_t)j * ld; };
    ED_TRY(ed_random_vector(n, seed, CPLX ? 1 : 0, vec(0), 1, st));

    const double eps23 = std::pow(2.220446049250313e-16, 2.0 / 3.0);
    std::vector<double> T((size_t)m * m, 0.0), theta, Z, c1(2 * m), c2(2 * m);
    std::vector<double> beta(m, 0.0);
    int k = 0, iter = 0, nconv = 0, m_eff = m;
    bool exhausted = false;
    double anorm = 0.0;
    for (;;) {
        // extend the Lanczos basis from column k to m
        m_eff = m;
        for (int j = k; j < m; ++j) {
            double *w = vec(j + 1);
            ED_TRY(hv_plain(*h, CPLX, vec(j), w, st));
            int64_t nbv;
            double b = 0.0;
            ED_TRY(proj_subtract<CPLX>(n, ld, j + 1, W.V, w, W, nb, c1.data(), st));
            ED_TRY(proj_subtract<CPLX>(n, ld, j + 1, W.V, w, W, nb, c2.data(), st));
            ED_TRY(norm2(nreal, w, W, &nbv, &b, st)); // synchronizes: c1, c2 are on the host now
            if (j == k) {
                for (int i = 0; i < k; ++i) T[(size_t)i * m + k] = T[(size_t)k * m + i] = c1[2 * i] + c2[2 * i];
            } else {
                T[(size_t)(j - 1) * m + j] = T[(size_t)j * m + j - 1] = beta[j - 1];
            }
            T[(size_t)j * m + j] = c1[2 * j] + c2[2 * j];
            beta[j] = b;
            anorm = std::max(anorm, std::fabs(T[(size_t)j * m + j]) + b);
            if (b <= 1e-13 * anorm || j + 1 == n) { // invariant subspace: T is exact on the Krylov space
                m_eff = j + 1;
                exhausted = true;
                break;
            }
            ED_TRY(ed_scal(n, CPLX ? 1 : 0, 1.0 / b, w, st));
        }
        std::vector<double> Tm((size_t)m_eff * m_eff);
        for (int r = 0; r < m_eff; ++r)
            for (int c = 0; c < m_eff; ++c) Tm[(size_t)r * m_eff + c] = T[(size_t)r * m + c];
        jacobi_eig(m_eff, Tm, theta, Z);
        const double bm = beta[m_eff - 1];
        nconv = 0;
        const int want = std::min(nev, m_eff);
        for (int i = 0; i < want; ++i) {
            const double res = exhausted ? 0.0 : std::fabs(bm * Z[(size_t)(m_eff - 1) * m_eff + i]);
            if (res <= tol * std::max(eps23, std::fabs(theta[i]))) ++nconv;
        }
        for (int i = 0; i < want; ++i) evals[i] = theta[i];
        ++iter;
        if (nconv >= want || iter >= maxit) break;
        // thick restart: keep the k lowest Ritz vectors, the residual direction becomes column k
        k = std::min(m_eff - 2, nev + (m_eff - nev) / 2);
        if (k < nev) k = std::min(nev, m_eff - 1);
        std::vector<double> Yk((size_t)m_eff * k);
        for (int j = 0; j < m_eff; ++j)
            for (int kk = 0; kk < k; ++kk) Yk[(size_t)j * k + kk] = Z[(size_t)j * m_eff + kk];
        ED_HIP(hipMemcpyAsync(W.Y, Yk.data(), sizeof(double) * Yk.size(), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL((k_restart<CPLX>), dim3((unsigned)((n + kEB - 1) / kEB)), dim3(kEB), 0, st, n, ld, m_eff,
                           k, W.V, (const double *)W.Y);
        ED_CHECK_LAUNCH();
        ED_HIP(hipMemcpyAsync(vec(k), vec(m_eff), sizeof(double) * (size_t)ld, hipMemcpyDeviceToDevice, st));
        std::fill(T.begin(), T.end(), 0.0);
        for (int i = 0; i < k; ++i) T[(size_t)i * m + i] = theta[i];
        ED_HIP(hipStreamSynchronize(st)); // Yk leaves scope
    }
    const int want = std::min(nev, m_eff);
    if (d_evecs) {
        std::vector<double> Yk((size_t)m_eff * want);
        for (int j = 0; j < m_eff; ++j)
            for (int kk = 0; kk < want; ++kk) Yk[(size_t)j * want + kk] = Z[(size_t)j * m_eff + kk];
        ED_HIP(hipMemcpyAsync(W.Y, Yk.data(), sizeof(double) * Yk.size(), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL((k_restart<CPLX>), dim3((unsigned)((n + kEB - 1) / kEB)), dim3(kEB), 0, st, n, ld, m_eff,
                           want, W.V, (const double *)W.Y);
        ED_CHECK_LAUNCH();
        ED_HIP(hipMemcpyAsync(d_evecs, W.V, sizeof(double) * (size_t)ld * want, hipMemcpyDeviceToDevice, st));
    }
    ED_HIP(hipStreamSynchronize(st));
    if (nconv_out) *nconv_out = nconv;
    if (niter_out) *niter_out = iter;
    return nconv >= want ? ED_OK : ED_ENOCONV;
}

} // namespace ed

using namespace ed;

extern "C" int ed_eigs_lowest(ed_hubbard hh, int complex_vectors, int nev, int ncv, int maxit, double tol,
                              uint32_t seed, double *evals, void *d_evecs, int *nconv, int *niter, void *stream)
{
    const Hubbard *h = reinterpret_cast<const Hubbard *>(hh);
    if (!h || !evals || nev < 1 || ncv < nev + 1 || ncv > kMaxNcv || maxit < 1 || !(tol > 0))
        return fail(ED_EINVAL, "ed_eigs_lowest: bad arguments (need 1 <= nev < ncv <= %d, maxit >= 1, tol > 0)",
                    kMaxNcv);
    const int64_t n = h->args.dim_up * h->args.dim_dn;
    if (nev > n) return fail(ED_EINVAL, "ed_eigs_lowest: nev = %d exceeds the dimension %lld", nev, (long long)n);
    if (!complex_vectors && h->cplx_tables)
        return fail(ED_EINVAL, "ed_eigs_lowest: complex tables need complex vectors");
    hipStream_t st = stream_of(stream);
    return complex_vectors ? eigs_run<true>(h, nev, ncv, maxit, tol, seed, evals, d_evecs, nconv, niter, st)
                           : eigs_run<false>(h, nev, ncv, maxit, tol, seed, evals, d_evecs, nconv, niter, st);
}

extern "C" int ed_hubbard_apply_host(ed_hubbard hh, int complex_vectors, const void *x, void *y, void *stream)
{
    const Hubbard *h = reinterpret_cast<const Hubbard *>(hh);
    if (!h || !x || !y) return fail(ED_EINVAL, "ed_hubbard_apply_host: bad arguments");
    const int64_t n = h->args.dim_up * h->args.dim_dn;
    const size_t bytes = sizeof(double) * (size_t)n * (complex_vectors ? 2 : 1);
    hipStream_t st = stream_of(stream);
    void *dx = nullptr, *dy = nullptr;
    ED_TRY(dmalloc(&dx, bytes));
    int s = dmalloc(&dy, bytes);
    if (!s && hipMemcpyAsync(dx, x, bytes, hipMemcpyHostToDevice, st) != hipSuccess)
        s = fail(ED_EHIP, "ed_hubbard_apply_host: upload failed");
    if (!s) s = ed_hubbard_apply(hh, complex_vectors, dx, dy, stream);
    if (!s && (hipMemcpyAsync(y, dy, bytes, hipMemcpyDeviceToHost, st) != hipSuccess ||
               hipStreamSynchronize(st) != hipSuccess))
        s = fail(ED_EHIP, "ed_hubbard_apply_host: download failed");
    if (st == nullptr || s) (void)hipDeviceSynchronize();
    dfree(dx);
    dfree(dy);
    return s;
}

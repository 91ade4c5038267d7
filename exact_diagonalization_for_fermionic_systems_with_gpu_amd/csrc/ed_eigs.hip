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
#include <cstdlib>
#include <vector>

#include "ed_common.hpp"
#include "ed_hubbard.hpp"

namespace ed {

constexpr int kEB = 256;
constexpr int kEChunk = kEB * 8;
constexpr int kMaxNcv = 64;

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

// c = V^T w over V_0..V_{m-1}: per block of kEChunk elements, w is read once into registers
// and each basis vector streamed against it; partial[j][block] = (Re, Im) of that block's sum
// of conj(V_j) w (the layout k_proj_fin reduces in a fixed order)
template <bool CPLX>
__global__ __launch_bounds__(kEB) void k_proj_all(int64_t n, int64_t ld, int m, const double *__restrict__ V,
                                                  const double *__restrict__ w, double *__restrict__ partial,
                                                  int64_t nb)
{
    constexpr int NC = CPLX ? 2 : 1, E = kEChunk / kEB;
    __shared__ double red[kEB / 64][kMaxNcv][2];
    const int64_t base = (int64_t)blockIdx.x * kEChunk + threadIdx.x;
    double wr[E], wi[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int64_t i = base + (int64_t)e * kEB;
        wr[e] = i < n ? w[NC * i] : 0.0;
        wi[e] = (CPLX && i < n) ? w[NC * i + 1] : 0.0;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int j = 0; j < m; ++j) {
        const double *v = V + (int64_t)j * ld;
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int64_t i = base + (int64_t)e * kEB;
            if (i < n) {
                if (CPLX) {
                    const double vr = v[2 * i], vi = v[2 * i + 1];
                    sr = fma(vr, wr[e], fma(vi, wi[e], sr));
                    si = fma(vr, wi[e], fma(-vi, wr[e], si));
                } else {
                    sr = fma(v[i], wr[e], sr);
                }
            }
        }
        sr = wave_sum(sr);
        if (CPLX) si = wave_sum(si);
        if (lane == 0) {
            red[wid][j][0] = sr;
            red[wid][j][1] = si;
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < m; j += kEB) {
        double sr = 0.0, si = 0.0;
        for (int q = 0; q < kEB / 64; ++q) {
            sr += red[q][j][0];
            si += red[q][j][1];
        }
        partial[2 * ((int64_t)j * nb + blockIdx.x)] = sr;
        partial[2 * ((int64_t)j * nb + blockIdx.x) + 1] = si;
    }
}

// w -= sum_{j<m} V_j c_j; NRM: also nrm[block] = that block's sum of |w|^2 after the update
template <bool CPLX, bool NRM>
__global__ __launch_bounds__(kEB) void k_sub_all(int64_t n, int64_t ld, int m, const double *__restrict__ V,
                                                 const double *__restrict__ c, double *__restrict__ w,
                                                 double *__restrict__ nrm)
{
    constexpr int NC = CPLX ? 2 : 1, E = kEChunk / kEB;
    __shared__ double nred[kEB / 64];
    const int64_t base = (int64_t)blockIdx.x * kEChunk + threadIdx.x;
    double wr[E], wi[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int64_t i = base + (int64_t)e * kEB;
        wr[e] = i < n ? w[NC * i] : 0.0;
        wi[e] = (CPLX && i < n) ? w[NC * i + 1] : 0.0;
    }
    for (int j = 0; j < m; ++j) {
        const double *v = V + (int64_t)j * ld;
        const double cr = c[2 * j], ci = c[2 * j + 1];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int64_t i = base + (int64_t)e * kEB;
            if (i < n) {
                if (CPLX) {
                    const double vr = v[2 * i], vi = v[2 * i + 1];
                    wr[e] -= vr * cr - vi * ci;
                    wi[e] -= vr * ci + vi * cr;
                } else {
                    wr[e] = fma(-cr, v[i], wr[e]);
                }
            }
        }
    }
    double t = 0.0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int64_t i = base + (int64_t)e * kEB;
        if (i < n) {
            w[NC * i] = wr[e];
            if (CPLX) w[NC * i + 1] = wi[e];
            if (NRM) t = fma(wr[e], wr[e], CPLX ? fma(wi[e], wi[e], t) : t);
        }
    }
    if (NRM) {
        t = wave_sum(t);
        if ((threadIdx.x & 63) == 0) nred[threadIdx.x >> 6] = t;
        __syncthreads();
        if (threadIdx.x == 0) {
            double r = 0.0;
            for (int q = 0; q < kEB / 64; ++q) r += nred[q];
            nrm[blockIdx.x] = r;
        }
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

// buffers of one solve, carved from the calling thread's arena (see Scratch)
struct EigWork {
    double *V = nullptr;   // (ncv + 1) vectors of ld doubles
    double *part = nullptr; // 2 * ncv * nb
    double *c = nullptr;    // 2 * ncv
    double *c2 = nullptr;   // 2 * ncv
    double *Y = nullptr;    // ncv * ncv
    double *nrm = nullptr;  // 1
};
static thread_local Arena t_eigs_arena;

// Classical Gram-Schmidt against V_0..V_{m-1}: c1 = V^T w, w -= V c1 with |w|^2 in the same
// pass; then, unless the first pass kept most of w (the DGKS test |w_new|^2 >= |w_old|^2 / 2,
// with |w_old|^2 = |c1|^2 + |w_new|^2; always when force2), c2 = V^T w, w -= V c2.  c1, c2 land
// in c1h, c2h (c2 = 0 after one pass) and |w| in *b; the call synchronises the stream.
template <bool CPLX>
static int cgs(bool force2, int64_t n, int64_t ld, int m, const double *V, double *w, EigWork &W, int64_t nb,
               double *c1h, double *c2h, double *b, hipStream_t st)
{
    const dim3 g((unsigned)nb), t(kEB);
    double *cc[2] = {W.c, W.c2}, *ch[2] = {c1h, c2h};
    std::fill(c2h, c2h + 2 * m, 0.0);
    for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL((k_proj_all<CPLX>), g, t, 0, st, n, ld, m, V, (const double *)w, W.part, nb);
        hipLaunchKernelGGL(k_proj_fin, dim3((unsigned)m), dim3(256), 0, st, (const double *)W.part, nb, cc[pass]);
        ED_CHECK_LAUNCH();
        ED_HIP(hipMemcpyAsync(ch[pass], cc[pass], sizeof(double) * 2 * m, hipMemcpyDeviceToHost, st));
        hipLaunchKernelGGL((k_sub_all<CPLX, true>), g, t, 0, st, n, ld, m, V, (const double *)cc[pass], w, W.part);
        ED_CHECK_LAUNCH();
        ED_TRY(finalize_sum(W.part, nb, W.nrm, 1, st));
        ED_HIP(hipMemcpyAsync(b, W.nrm, sizeof(double), hipMemcpyDeviceToHost, st));
        ED_HIP(hipStreamSynchronize(st));
        if (pass == 0 && !force2) {
            double c2sum = 0.0;
            for (int i = 0; i < 2 * m; ++i) c2sum += c1h[i] * c1h[i];
            if ((*b) * (*b) >= 0.5 * (c2sum + (*b) * (*b))) break;
        }
    }
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
    Scratch S(t_eigs_arena, st);
    const size_t bV = sizeof(double) * (size_t)ld * (m + 1),
                 bP = sizeof(double) * 2 * (size_t)std::max<int64_t>(m * nb, nbblas), bC = sizeof(double) * 2 * m,
                 bY = sizeof(double) * m * m, bN = sizeof(double);
    ED_TRY(S.reserve(Scratch::round(bV) + Scratch::round(bP) + 2 * Scratch::round(bC) + Scratch::round(bY) +
                     Scratch::round(bN)));
    W.V = (double *)S.take(bV);
    W.part = (double *)S.take(bP);
    W.c = (double *)S.take(bC);
    W.c2 = (double *)S.take(bC);
    W.Y = (double *)S.take(bY);
    W.nrm = (double *)S.take(bN);
    auto vec = [&](int j) { return W.V + (int64_t)j * ld; };
    ED_TRY(ed_random_vector(n, seed, CPLX ? 1 : 0, vec(0), 1, st));

    const double eps23 = std::pow(2.220446049250313e-16, 2.0 / 3.0);
    const char *ep = getenv("ED_EIGS_PASSES");
    const bool two_pass = ep && atoi(ep) == 2;
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
            double b = 0.0;
            // second pass by the DGKS test (ED_EIGS_PASSES=2: always)
            ED_TRY(cgs<CPLX>(two_pass, n, ld, j + 1, W.V, w, W, nb, c1.data(), c2.data(), &b, st));
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
    static thread_local Arena t_apply_arena;
    Scratch S(t_apply_arena, st);
    ED_TRY(S.reserve(2 * Scratch::round(bytes)));
    void *dx = S.take(bytes), *dy = S.take(bytes);
    int s = ED_OK;
    if (hipMemcpyAsync(dx, x, bytes, hipMemcpyHostToDevice, st) != hipSuccess)
        s = fail(ED_EHIP, "ed_hubbard_apply_host: upload failed");
    if (!s) s = ed_hubbard_apply(hh, complex_vectors, dx, dy, stream);
    if (!s && (hipMemcpyAsync(y, dy, bytes, hipMemcpyDeviceToHost, st) != hipSuccess ||
               hipStreamSynchronize(st) != hipSuccess))
        s = fail(ED_EHIP, "ed_hubbard_apply_host: download failed");
    if (st == nullptr || s) (void)hipDeviceSynchronize();
    return s;
}

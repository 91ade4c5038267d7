// ed_lanczos.hip -- host driver of the FP64 Lanczos ground-state solver and the
// fixed-length tridiagonalisation used by the continued fraction.
//
// Reference: Lanczos<T>::eigen / tridiagonal (MBL_test/include/HubMatrix_GPU.cuh:541-613;
// complex twin in TotalSpin/include/LanczosSolver.cuh:111-238).  The recurrence, the
// anorm update, the N%10==9 convergence test |b_{N+1} z_N| < tol*anorm && |z_0| > tol*anorm
// and the Ritz vector rebuilt by re-running the recurrence from the saved start are the
// reference's; what changes is where the work runs:
//   * one Lanczos step is two kernels: the H.v with the 1/b scaling, the -b v_{N-1}
//     term and the <v_N, w> partials fused in (ed_hubbard.hip), then f = w - a v_N fused
//     with the |f|^2 partials (ed_blas.hip).  Three resident vectors rotate roles.
//   * a and b live on the device; the host reads them only at the convergence checks,
//     so steps between checks are enqueued without a host round trip.
//   * the tridiagonal eigenproblem (LAPACK dstedc in the reference) is solved on the
//     host by Sturm bisection for the lowest eigenvalue and inverse iteration for its
//     vector; all eigenvalues by the implicit QL method.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "ed_common.hpp"
#include "ed_hubbard.hpp"
#include "ed_thread.hpp"
#include "ed_tridiag.hpp"

namespace ed {

// ---- symmetric tridiagonal eigen-solvers (host) ------------------------------------

// the lowest eigenpair: Sturm bisection and inverse iteration (ed_tridiag.hpp, shared with the
// batch engine's device-side convergence test)
static double bisect_kth(int n, const double *d, const double *e, int k) { return tri_bisect_kth(n, d, e, k); }

static void inverse_iteration(int n, const double *d, const double *e, double lam, double *z)
{
    std::vector<double> ws(5 * (size_t)n);
    std::vector<char> sw((size_t)n);
    tri_inverse_iteration(n, d, e, lam, z, ws.data(), sw.data());
}

// implicit QL with Wilkinson-type shifts, Golub & Van Loan sec. 8.3; with z (n x n, column j =
// eigenvector j, z[j * n + i], identity on entry) the plane rotations are accumulated into it
static void ql_eig(int n, double *d, double *e /* n entries, e[n-1] scratch */, double *z)
{
    e[n - 1] = 0.0;
    for (int l = 0; l < n; ++l) {
        int iter = 0, m;
        do {
            for (m = l; m < n - 1; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= std::numeric_limits<double>::epsilon() * dd) break;
            }
            if (m != l) {
                if (++iter > 60) break;
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = std::hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
                double s = 1.0, c = 1.0, p = 0.0;
                int i;
                for (i = m - 1; i >= l; --i) {
                    double f = s * e[i], b = c * e[i];
                    r = std::hypot(f, g);
                    e[i + 1] = r;
                    if (r == 0.0) {
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    p = s * r;
                    d[i + 1] = g + p;
                    g = c * r - b;
                    if (z) {
                        double *zi = z + (size_t)i * n, *zj = z + (size_t)(i + 1) * n;
                        for (int k = 0; k < n; ++k) {
                            const double t = zj[k];
                            zj[k] = s * zi[k] + c * t;
                            zi[k] = c * zi[k] - s * t;
                        }
                    }
                }
                if (r == 0.0 && i >= l) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
}

// ---- orthogonality monitor ---------------------------------------------------------
// Estimates w_{j,k} = v_j . v_k of the Lanczos vectors from a and b alone (the omega recurrence
// of H. D. Simon, "The Lanczos algorithm with partial reorthogonalization", Math. Comp. 42 (1984)
// 115-142).  Taking v_k . (H v_j) - v_j . (H v_k) = 0 with H v_j = b[j] v_{j-1} + a[j] v_j +
// b[j+1] v_{j+1} gives, for k < j,
//   b[j+1] w_{j+1,k} = b[k+1] w_{j,k+1} + (a[k] - a[j]) w_{j,k} + b[k] w_{j,k-1} - b[j] w_{j-1,k}
// plus the rounding of the two steps, modelled as in that paper by eps (b[k+1] + b[j+1]) times a
// normal deviate of standard deviation 0.3 (a fixed pseudo-random sequence, so the estimate is
// reproducible); w_{j+1,j} = eps |T| / b[j+1] with |T| the running max_k |a[k]| + b[k] + b[k+1],
// and w_{j,j} = 1.
struct OmegaMonitor {
    std::vector<double> prev, cur{1.0}, next;
    int j = 0; // cur holds row j
    double tnorm = 0.0;
    uint64_t rng = 0x9e3779b97f4a7c15ull;
    double normal() // Box-Muller over splitmix64 uniforms
    {
        auto u01 = [&]() {
            uint64_t z = (rng += 0x9e3779b97f4a7c15ull);
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            z ^= z >> 31;
            return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        };
        const double u = u01(), v = u01();
        return std::sqrt(-2.0 * std::log(u)) * std::cos(6.283185307179586 * v);
    }
    // row j+1 from a[0..j], b[0..j+1]; returns max_{k<=j} |w_{j+1,k}|
    double step(const double *a, const double *b)
    {
        const double eps = std::numeric_limits<double>::epsilon();
        tnorm = std::fmax(tnorm, std::fabs(a[j]) + (j > 0 ? b[j] : 0.0) + b[j + 1]);
        const double bj1 = b[j + 1] > 0.0 ? b[j + 1] : std::numeric_limits<double>::min();
        next.assign((size_t)j + 2, 0.0);
        double mx = 0.0;
        for (int k = 0; k < j; ++k) {
            double t = b[k + 1] * cur[k + 1] + (a[k] - a[j]) * cur[k];
            if (k > 0) t += b[k] * cur[k - 1];
            t -= b[j] * prev[k];
            t += 0.3 * eps * (b[k + 1] + b[j + 1]) * normal();
            next[k] = t / bj1;
            mx = std::fmax(mx, std::fabs(next[k]));
        }
        next[j] = eps * tnorm / bj1;
        next[j + 1] = 1.0;
        mx = std::fmax(mx, next[j]);
        prev.swap(cur);
        cur.swap(next);
        ++j;
        return mx;
    }
    // after the two newest vectors were orthogonalised against every direction the loss lies in
    void reset_recent()
    {
        const double eps = std::numeric_limits<double>::epsilon();
        for (int k = 0; k + 1 < j; ++k) cur[k] = eps;
        for (int k = 0; k + 1 < j - 1; ++k) prev[k] = eps;
    }
};

// ---- Lanczos workspace -------------------------------------------------------------
struct Work {
    double *buf[2] = {nullptr, nullptr};
    double *partial = nullptr;
    double *adev = nullptr, *bdev = nullptr, *ddev = nullptr;
    ~Work()
    {
        dfree(buf[0]);
        dfree(buf[1]);
        dfree(partial);
        dfree(adev);
        dfree(bdev);
        dfree(ddev);
    }
};

static int alloc_work(Work &w, const Hubbard &h, int64_t nreal, int maxN, bool vecs)
{
    if (vecs) {
        ED_TRY(dmalloc((void **)&w.buf[0], sizeof(double) * (size_t)nreal));
        ED_TRY(dmalloc((void **)&w.buf[1], sizeof(double) * (size_t)nreal));
    }
    const int64_t np = std::max<int64_t>(std::max<int64_t>(hv_num_blocks(h), slab_num_partials(h)),
                                         blas_num_blocks(nreal));
    ED_TRY(dmalloc((void **)&w.partial, sizeof(double) * (size_t)np));
    ED_TRY(dmalloc((void **)&w.adev, sizeof(double) * (size_t)(maxN + 2)));
    ED_TRY(dmalloc((void **)&w.bdev, sizeof(double) * (size_t)(maxN + 3)));
    ED_TRY(dmalloc((void **)&w.ddev, sizeof(double) * (size_t)(maxN + 2)));
    return ED_OK;
}

// One step N: (A = f, B = v_{N-1}) -> (C = v_N, B = f_{N+1}); a[N], b[N+1] on the device.
static int lanczos_step(const Hubbard &h, bool cvec, int64_t nreal, double *A, double *B, double *C, Work &w,
                        int N, bool reorth, bool slab, hipStream_t st)
{
    int64_t nb;
    if (slab) ED_TRY(slab_lanczos(h, A, B, C, B, w.bdev, N, true, w.partial, st, &nb));
    else ED_TRY(hv_lanczos(h, cvec, A, B, C, B, w.bdev, N, w.partial, st, &nb));
    ED_TRY(finalize_sum(w.partial, nb, w.adev + N, 0, st));
    ED_TRY(lanczos_update(nreal, C, B, w.adev, N, w.partial, &nb, st));
    if (reorth) {
        // second Gram-Schmidt pass against v_N (Daniel-Gragg-Kaufman-Stewart correction):
        // delta = <v_N, f>; f -= delta v_N; the tridiagonal entry is a[N] + delta
        ED_TRY(dot_partials(nreal, C, B, w.partial, &nb, st));
        ED_TRY(finalize_sum(w.partial, nb, w.ddev + N, 0, st));
        ED_TRY(lanczos_update(nreal, C, B, w.ddev, N, w.partial, &nb, st));
    }
    ED_TRY(finalize_sum(w.partial, nb, w.bdev + N + 1, 1, st));
    return ED_OK;
}

// One step on the slab layout without v: (A = f_N, B = f_{N-1}) -> C = f_{N+1}, a[N], b[N+1] on the
// device.  v_N = f_N / b[N] is formed where it is read (the H.v pass and the update), exactly as
// the step with a stored v forms it, so a, b and the vectors are the same bit for bit; the step
// writes one vector less.  Caller's rotation: (A, B, C) <- (C, A, B).
static int lanczos_step_vfree(const Hubbard &h, int64_t nreal, double *A, double *B, double *C, Work &w, int N,
                              hipStream_t st)
{
    int64_t nb;
    ED_TRY(slab_lanczos_vfree(h, A, B, C, w.bdev, N, w.partial, st, &nb));
    ED_TRY(finalize_sum(w.partial, nb, w.adev + N, 0, st));
    ED_TRY(lanczos_update_vfree(nreal, A, C, w.adev, w.bdev, N, w.partial, &nb, st));
    ED_TRY(finalize_sum(w.partial, nb, w.bdev + N + 1, 1, st));
    return ED_OK;
}

// The nev lowest eigenvalues of a Lanczos tridiagonal T (a[0..M-1], couplings e[0..M-2]), after
// the test of J. Cullum and R. A. Willoughby ("Lanczos Algorithms for Large Symmetric Eigenvalue
// Computations", vol. 1, ch. 4).  Eigenvalues of T within tol_m of each other form a cluster: two
// or more copies are a converged eigenvalue of H (a copy appears only once the eigenvalue has
// converged) and count once.  A simple eigenvalue that is not also an eigenvalue of T^ (T with its
// first row and column removed) stands for an eigenvalue of H and must pass |bres s[M-1]| <
// tol * anorm (s its unit eigenvector of T, bres the coupling after the last step), the residual
// test of ed_lanczos_eigen.  A simple eigenvalue that is also one of T^ has a tiny first component
// s[0]: a ghost, or an eigenvalue of H that the start vector barely touches (an overlap of 1e-7
// puts it within 1e-12 |T| of T^); in a disordered 14-site chain a ground state 1e-7 from the
// start vector converged after its excited levels.
//   * after a recurrence without reorthogonalisation: converged, it holds the test back until a
//     copy of it forms and merges (or max_steps ends the run); not converged, it is a ghost in
//     transit and dropped when it lies within 1e-3 |T| above a converged value kept below it (a
//     ghost converges from above onto its twin), and otherwise holds the test back: it may be an
//     eigenvalue coming in, and T's lowest eigenvalue cannot be a ghost at all;
//   * after selective orthogonalisation (semi_orth), which forms ghosts only around eigenvalues
//     closer together than its accuracy: converged, a ghost when the converged value just below
//     it lies within sqrt(tol) |T|, otherwise kept; not converged, it holds the test back.
// Left open: an eigenvalue the start vector barely touches that lies within 1e-3 |T| above
// another one (1e-7 apart in a test) is taken for a ghost until it converges.  Returns the number
// found (at most nev); *conv = every one found passes.
int cw_lowest(int M, const double *a, const double *e, double bres, int nev, double tol, double anorm,
              bool semi_orth, double *out, bool *conv)
{
    *conv = false;
    if (M < 1) return 0;
    std::vector<double> th((size_t)M), thh((size_t)std::max(M - 1, 1));
    ed_tridiag_eigvals(M, a, e, th.data());
    if (M > 1) ed_tridiag_eigvals(M - 1, a + 1, e + 1, thh.data());
    // multiplicity and spurious-value tolerance: copies of a converged eigenvalue agree to a few
    // hundred eps |T|, while distinct eigenvalues of H can lie 1e-8 apart
    const double tm = 1e-12 * anorm;
    const double tg = std::sqrt(tol) * anorm; // a ghost's distance to its twin while it converges
    std::vector<double> ws(5 * (size_t)M), z((size_t)M);
    std::vector<char> sw((size_t)M);
    int found = 0;
    bool ok = true;
    // the previous cluster: its largest member, and whether it stands for a converged eigenvalue
    double prev_hi = -std::numeric_limits<double>::infinity();
    bool prev_genuine = false;
    // the largest converged value kept so far, and the window above it in which an unconverged
    // value on T^ is taken for a ghost approaching it
    double genuine_hi = -std::numeric_limits<double>::infinity();
    const double tw = 1e-3 * anorm;
    for (int j = 0; j < M && found < nev;) {
        int r = j + 1;
        while (r < M && th[r] - th[r - 1] <= tm) ++r;
        const double v = th[j];
        bool keep = true, genuine = true;
        if (r - j == 1) {
            tri_inverse_iteration(M, a, e, v, z.data(), ws.data(), sw.data());
            bool converged = std::fabs(bres * z[M - 1]) < tol * anorm;
            bool in_that = false;
            if (M > 1) {
                const double *q = std::lower_bound(thh.data(), thh.data() + (M - 1), v - tm);
                in_that = q < thh.data() + (M - 1) && *q <= v + tm;
            }
            if (in_that) {
                if (converged && !semi_orth) converged = false; // held back until its copy forms
                else if (converged) keep = !(prev_genuine && v - prev_hi <= tg);
                else keep = semi_orth || v - genuine_hi > tw;
            }
            if (keep) ok = ok && converged;
            genuine = keep && converged;
        }
        if (keep) out[found++] = v;
        prev_hi = th[r - 1];
        prev_genuine = genuine;
        if (genuine) genuine_hi = th[r - 1];
        j = r;
    }
    *conv = ok && found == nev;
    return found;
}

// ---- selective orthogonalisation (Parlett & Scott) -----------------------------------
// B. N. Parlett and D. S. Scott, "The Lanczos algorithm with selective orthogonalization",
// Math. Comp. 33 (1979) 217-238.  The Lanczos vectors lose orthogonality only along Ritz vectors
// that have converged ("good" ones: b_{N+1} |s_N| <= sqrt(eps) |T|, Paige's theorem), and once a
// Ritz vector is good the loss along it regrows within a step, so every new f is orthogonalised
// against all good Ritz vectors (on the device, no host round trip).  The Ritz analysis that finds
// new good vectors runs at the convergence-test cadence; a vector found good there is also
// removed from v_N.  Forming a Ritz vector needs the Lanczos basis, which is kept on the device;
// the Ritz vector of the result is then V z (no second recurrence).

__global__ __launch_bounds__(256) void k_basis_combine(int64_t n, int nk, const double *__restrict__ V,
                                                       const double *__restrict__ s, double *__restrict__ y,
                                                       int accumulate)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double acc = accumulate ? y[i] : 0.0;
    for (int k = 0; k < nk; ++k) acc = fma(s[k], V[(size_t)k * (size_t)n + (size_t)i], acc);
    y[i] = acc;
}

// i y of an interleaved complex vector
__global__ __launch_bounds__(256) void k_times_i(int64_t ncplx, const double *__restrict__ y, double *__restrict__ iy)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= ncplx) return;
    iy[2 * i] = -y[2 * i + 1];
    iy[2 * i + 1] = y[2 * i];
}

// The basis chunks and good Ritz vectors come from a grow-only pool of the calling thread, reused by
// its next solve (a solve with a few hundred vectors would otherwise map and unmap gigabytes per
// call); released when the thread exits (not during process teardown, ed_thread.hpp), or between
// solves when a device allocation of the thread fails (release_idle_thread_pools).
struct SoPool {
    std::vector<double *> chunk; // basis chunks
    std::vector<size_t> chunk_cap;
    std::vector<double *> yv;    // Ritz vectors
    std::vector<size_t> yv_cap;
    int users = 0;               // solves of this thread that have taken buffers of the pool
    size_t release()
    {
        size_t bytes = 0;
        for (size_t i = 0; i < chunk.size(); ++i) {
            dfree(chunk[i]);
            bytes += chunk_cap[i] * sizeof(double);
        }
        for (size_t i = 0; i < yv.size(); ++i) {
            dfree(yv[i]);
            bytes += yv_cap[i] * sizeof(double);
        }
        chunk.clear();
        chunk_cap.clear();
        yv.clear();
        yv_cap.clear();
        return bytes;
    }
    ~SoPool()
    {
        if (process_exiting()) return;
        release();
    }
    // the i-th buffer of list v (capacities cap) with room for `doubles` doubles
    static int get(std::vector<double *> &v, std::vector<size_t> &cap, size_t i, size_t doubles, double **out)
    {
        if (v.size() <= i) {
            v.resize(i + 1, nullptr);
            cap.resize(i + 1, 0);
        }
        if (cap[i] < doubles) {
            dfree(v[i]);
            v[i] = nullptr;
            cap[i] = 0;
            ED_TRY(dmalloc((void **)&v[i], sizeof(double) * doubles));
            cap[i] = doubles;
        }
        *out = v[i];
        return ED_OK;
    }
};
static thread_local SoPool t_so_pool;

size_t release_idle_thread_pools() { return t_so_pool.users == 0 ? t_so_pool.release() : 0; }

struct SelOrth {
    static constexpr int kChunk = 16;
    int64_t n = 0; // doubles per vector
    bool cvec = false;
    std::vector<double *> chunks; // v_0, v_1, ... kChunk per allocation (t_so_pool's)
    std::vector<double *> Y;      // unit good Ritz vectors (with i y after each, complex vectors)
    std::vector<double> theta;    // their Ritz values
    double *coef = nullptr;       // device: purge coefficients, then Ritz coordinates
    int coef_cap = 0;
    // Ritz values above theta_cap are not kept (the nev-lowest solve sets it to its nev-th lowest
    // Ritz value: copies of converged eigenvalues further up stay up there and do not disturb the
    // lowest ones, so the basis need only stay orthogonal to the converged vectors at the low end)
    double theta_cap = std::numeric_limits<double>::infinity();
    int analyses = 0, purges = 0;
    double *part = nullptr;       // device: dot partials of the duplicate check
    bool holds = false;           // counted in t_so_pool.users (from its first take on)
    SelOrth() = default;
    SelOrth(const SelOrth &) = delete;
    SelOrth &operator=(const SelOrth &) = delete;
    ~SelOrth()
    {
        if (holds) --t_so_pool.users;
        dfree(coef);
        dfree(part);
    }
    // before the first take: the pool is in use from here on (a failed allocation meanwhile,
    // its own included, must not release it)
    void hold()
    {
        if (!holds) {
            holds = true;
            ++t_so_pool.users;
        }
    }
    double *vec(int k) { return chunks[(size_t)k / kChunk] + (size_t)(k % kChunk) * (size_t)n; }
    int new_y(double **y)
    {
        hold();
        return SoPool::get(t_so_pool.yv, t_so_pool.yv_cap, Y.size(), (size_t)n, y);
    }
    int store(int k, const double *v, hipStream_t st)
    {
        hold();
        while ((int)chunks.size() * kChunk <= k) {
            double *c = nullptr;
            if (SoPool::get(t_so_pool.chunk, t_so_pool.chunk_cap, chunks.size(), (size_t)n * kChunk, &c) != ED_OK)
                return fail(ED_ENOMEM, "ed_lanczos_eigen: selective reorthogonalisation keeps the Lanczos basis on "
                                       "the device; %d vectors of %lld doubles do not fit",
                            (int)(chunks.size() + 1) * kChunk, (long long)n);
            chunks.push_back(c);
        }
        ED_HIP(hipMemcpyAsync(vec(k), v, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, st));
        return ED_OK;
    }
    int reserve_coef(int m)
    {
        if (m <= coef_cap) return ED_OK;
        dfree(coef);
        coef = nullptr;
        coef_cap = 0;
        ED_TRY(dmalloc((void **)&coef, sizeof(double) * (size_t)m));
        coef_cap = m;
        return ED_OK;
    }
    // y = sum_{k<=M} z_k v_k
    int combine(int M, const double *z, double *y, hipStream_t st)
    {
        ED_TRY(reserve_coef(std::max(M + 1, 2 * (int)Y.size() + 2)));
        ED_HIP(hipMemcpyAsync(coef, z, sizeof(double) * (size_t)(M + 1), hipMemcpyHostToDevice, st));
        const unsigned grid = (unsigned)((n + 255) / 256);
        for (int k0 = 0; k0 <= M; k0 += kChunk) {
            const int nk = std::min(kChunk, M + 1 - k0);
            hipLaunchKernelGGL(k_basis_combine, dim3(grid), dim3(256), 0, st, n, nk, (const double *)vec(k0),
                               (const double *)(coef + k0), y, k0 > 0 ? 1 : 0);
            ED_CHECK_LAUNCH();
        }
        // the host copy of z must outlive the upload
        ED_HIP(hipStreamSynchronize(st));
        return ED_OK;
    }
    // x -= (y . x) y for every good Ritz vector in turn; the partials of |x|^2 of the result
    // (nb = 0 when there is none)
    int purge(double *x, double *partial, int64_t *nb, hipStream_t st)
    {
        *nb = 0;
        ED_TRY(reserve_coef(2 * (int)Y.size() + 2));
        for (size_t q = 0; q < Y.size(); ++q) {
            int64_t nd;
            ED_TRY(dot_partials(n, Y[q], x, partial, &nd, st));
            ED_TRY(finalize_sum(partial, nd, coef + q, 0, st));
            ED_TRY(lanczos_update(n, Y[q], x, coef, (int)q, partial, nb, st));
        }
        return ED_OK;
    }
    // Ritz analysis of T (a[0..M], b[1..M]) with residual norm b[M+1]: good pairs not yet kept are
    // formed from the basis and kept; *found = number of new ones
    int analyse(int M, const double *a, const double *b, hipStream_t st, int *found)
    {
        const size_t before = Y.size();
        ++analyses;
        const double eps = std::numeric_limits<double>::epsilon();
        double tn = 1e-300;
        for (int k = 0; k <= M; ++k) tn = std::fmax(tn, std::fabs(a[k]) + (k > 0 ? b[k] : 0.0) + b[k + 1]);
        std::vector<double> d(a, a + M + 1), e((size_t)M + 1, 0.0), z((size_t)M + 1);
        for (int i = 0; i < M; ++i) e[i] = b[i + 1];
        ql_eig(M + 1, d.data(), e.data(), nullptr);
        std::sort(d.begin(), d.end());
        const int per = cvec ? 2 : 1; // entries of Y per kept Ritz vector (y, and i y for complex)
        for (int i = 0; i <= M; ++i) {
            const double th = d[i];
            if (th > theta_cap) break; // ascending
            // a kept vector's own Ritz value is found again here, moved by at most rounding: skip
            // it; one a little further away may still be the same vector (its value moves while it
            // converges) or a distinct eigenvalue close by (1e-8 apart happens), which the
            // overlap test below tells apart
            bool known = false;
            for (double t : theta) known = known || std::fabs(t - th) <= 1e-12 * tn;
            if (known) continue;
            inverse_iteration(M + 1, a, b + 1, th, z.data());
            if (!(std::fabs(b[M + 1] * z[M]) <= std::sqrt(eps) * tn)) continue;
            double *y = nullptr;
            ED_TRY(new_y(&y));
            ED_TRY(combine(M, z.data(), y, st));
            ED_TRY(ed_normalize(n, 0, y, nullptr, st));
            const int nkept = (int)theta.size() / per;
            int dup = -1;
            if (nkept > 0) {
                // |<y_q, y>| for every kept y_q (complex: from the real dots with y_q and i y_q)
                const int64_t nb = blas_num_blocks(n);
                if (!part) ED_TRY(dmalloc((void **)&part, sizeof(double) * (size_t)nb));
                ED_TRY(reserve_coef(std::max(M + 1, (int)Y.size() + 2)));
                for (size_t q = 0; q < Y.size(); ++q) {
                    int64_t nd;
                    ED_TRY(dot_partials(n, Y[q], y, part, &nd, st));
                    ED_TRY(finalize_sum(part, nd, coef + q, 0, st));
                }
                std::vector<double> ov(Y.size());
                ED_HIP(hipMemcpyAsync(ov.data(), coef, sizeof(double) * Y.size(), hipMemcpyDeviceToHost, st));
                ED_HIP(hipStreamSynchronize(st));
                for (int q = 0; q < nkept; ++q) {
                    double o2 = 0.0;
                    for (int c = 0; c < per; ++c) o2 += ov[(size_t)q * per + c] * ov[(size_t)q * per + c];
                    if (o2 > 0.25) dup = q;
                }
            }
            if (dup >= 0) { // the same vector again: remember its current value
                for (int c = 0; c < per; ++c) theta[(size_t)dup * per + c] = th;
                continue;
            }
            Y.push_back(y);
            theta.push_back(th);
            if (cvec) {
                double *iy = nullptr;
                ED_TRY(new_y(&iy));
                Y.push_back(iy);
                theta.push_back(th);
                hipLaunchKernelGGL(k_times_i, dim3((unsigned)((n / 2 + 255) / 256)), dim3(256), 0, st, n / 2,
                                   (const double *)y, iy);
                ED_CHECK_LAUNCH();
            }
        }
        *found = (int)(Y.size() - before);
        return ED_OK;
    }
};

// statistics of the calling thread's last ed_lanczos_eigen with selective reorthogonalisation
struct SoStats {
    int good = 0, analyses = 0, purges = 0;
    double omega_max = 0.0;
};
static thread_local SoStats t_so_stats;

} // namespace ed

using namespace ed;

extern "C" int ed_lanczos_so_stats(int *good, int *analyses, int *purges, double *omega_max)
{
    if (good) *good = t_so_stats.good;
    if (analyses) *analyses = t_so_stats.analyses;
    if (purges) *purges = t_so_stats.purges;
    if (omega_max) *omega_max = t_so_stats.omega_max;
    return ED_OK;
}

extern "C" int ed_lanczos_orthogonality(int n, const double *a, const double *b, double *omega_max)
{
    if (n < 1 || !a || !b || !omega_max) return fail(ED_EINVAL, "ed_lanczos_orthogonality: bad arguments");
    OmegaMonitor om;
    omega_max[0] = 0.0;
    for (int j = 1; j < n; ++j) omega_max[j] = om.step(a, b);
    return ED_OK;
}

extern "C" int ed_tridiag_lowest(int n, const double *d, const double *e, double *lambda, double *z)
{
    if (n < 1 || !d || (n > 1 && !e) || !lambda) return fail(ED_EINVAL, "ed_tridiag_lowest: bad arguments");
    const double lam = bisect_kth(n, d, e, 0);
    *lambda = lam;
    if (z) inverse_iteration(n, d, e, lam, z);
    return ED_OK;
}

extern "C" int ed_tridiag_eigvals(int n, const double *d, const double *e, double *w)
{
    if (n < 1 || !d || (n > 1 && !e) || !w) return fail(ED_EINVAL, "ed_tridiag_eigvals: bad arguments");
    std::vector<double> dd(d, d + n), ee(n, 0.0);
    for (int i = 0; i < n - 1; ++i) ee[i] = e[i];
    ql_eig(n, dd.data(), ee.data(), nullptr);
    std::sort(dd.begin(), dd.end());
    for (int i = 0; i < n; ++i) w[i] = dd[i];
    return ED_OK;
}

extern "C" int ed_tridiag_lanczos_lowest(int n, const double *d, const double *e, double bres, int nev, double tol,
                                         double anorm, int semi_orth, double *out, int *found, int *conv)
{
    if (n < 1 || !d || (n > 1 && !e) || nev < 1 || !out || !(tol > 0.0) || !(anorm > 0.0))
        return fail(ED_EINVAL, "ed_tridiag_lanczos_lowest: bad arguments");
    for (int k = 0; k < nev; ++k) out[k] = std::numeric_limits<double>::quiet_NaN();
    bool c = false;
    const int f = cw_lowest(n, d, e, bres, nev, tol, anorm, semi_orth != 0, out, &c);
    if (found) *found = f;
    if (conv) *conv = c ? 1 : 0;
    return ED_OK;
}

extern "C" int ed_tridiag_eigen(int n, const double *d, const double *e, double *w, double *z)
{
    if (n < 1 || !d || (n > 1 && !e) || !w || !z) return fail(ED_EINVAL, "ed_tridiag_eigen: bad arguments");
    std::vector<double> dd(d, d + n), ee(n, 0.0), zz((size_t)n * n, 0.0);
    for (int i = 0; i < n - 1; ++i) ee[i] = e[i];
    for (int i = 0; i < n; ++i) zz[(size_t)i * n + i] = 1.0;
    ql_eig(n, dd.data(), ee.data(), zz.data());
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return dd[x] < dd[y]; });
    for (int j = 0; j < n; ++j) {
        w[j] = dd[idx[j]];
        std::copy(zz.begin() + (size_t)idx[j] * n, zz.begin() + (size_t)(idx[j] + 1) * n, z + (size_t)j * n);
    }
    return ED_OK;
}

// The driver of ed_lanczos_eigen (nev = 1) and ed_lanczos_lowest (nev >= 1, selective
// orthogonalisation): eval[0..nev-1] get the nev lowest Ritz values of the last test, the Ritz
// vector is the lowest one's.
static int lanczos_eigen_impl(const Hubbard *h, int complex_vectors, const ed_lanczos_cfg *cfg, void *d_f, int nev,
                              double *eval, void *d_evec, double *a_out, double *b_out, int *nsteps, void *stream)
{
    if (!h || !d_f || !eval || d_f == d_evec || nev < 1) return fail(ED_EINVAL, "ed_lanczos_eigen: bad arguments");
    const bool want_vec = d_evec != nullptr;
    const bool cvec = complex_vectors != 0;
    if (h->cplx_tables && !cvec) return fail(ED_EINVAL, "ed_lanczos_eigen: complex tables need complex vectors");
    const int maxN = cfg && cfg->max_steps > 0 ? cfg->max_steps : 1000;
    const double tol = cfg && cfg->tol > 0 ? cfg->tol : 0.01 * std::sqrt(std::numeric_limits<double>::epsilon());
    const int every = cfg && cfg->check_every > 0 ? cfg->check_every : 10;
    const int rmode = nev > 1 ? 2 : (cfg ? cfg->reorth : 0);
    if (rmode < 0 || rmode > 2) return fail(ED_EINVAL, "ed_lanczos_eigen: reorth must be 0, 1 or 2");
    const bool reorth = rmode == 1, selective = rmode == 2;
    hipStream_t st = stream_of(stream);
    const int64_t n = h->args.dim_up * h->args.dim_dn;
    t_so_stats = SoStats();
    // real sectors with a slab layout (ed_hubbard_slab_layout, or built here for sectors of 2^27
    // amplitudes and more) run the whole recurrence on it: the start vector is converted once and
    // the Ritz vector converted back at the end
    // layout vectors the slab path allocates: w.buf[0..1], plus the copy of f and (with a Ritz
    // vector) of evec unless the layout has no padding and the caller's f and evec carry them
    if (!cvec) slab_auto(*const_cast<Hubbard *>(h), want_vec ? (slab_len_for(*h) == n ? 2 : 4) : 3);
    const bool slab = h->slab && !cvec;
    const int64_t nreal = slab ? slab_len(*h) : (cvec ? 2 * n : n);

    Work w;
    ED_TRY(alloc_work(w, *h, nreal, maxN, true));
    struct Own {
        double *p[2] = {nullptr, nullptr};
        ~Own()
        {
            dfree(p[0]);
            dfree(p[1]);
        }
    } own;
    double *F = (double *)d_f, *ev = (double *)d_evec;
    // On a layout without padding (nreal == n) the caller's vectors carry layout vectors: evec
    // holds the layout copy of f during the forward recurrence while f keeps the start vector
    // (reference layout) for the rebuild, which then accumulates the Ritz vector in f's buffer.
    // The solve then holds five vectors: f, evec, the two work vectors and the scratch copy.
    const bool inplace = slab && want_vec && nreal == n && !selective;
    if (inplace) {
        ED_TRY(slab_from_rowmajor(*h, (const double *)d_f, (double *)d_evec, st));
        F = (double *)d_evec;
    } else if (slab) {
        // layout-L copies of f and evec; padding entries stay zero in every vector
        ED_TRY(dmalloc((void **)&own.p[0], sizeof(double) * (size_t)nreal));
        if (want_vec) ED_TRY(dmalloc((void **)&own.p[1], sizeof(double) * (size_t)nreal));
        for (double *b : {w.buf[0], w.buf[1], own.p[0], own.p[1]})
            if (b) ED_HIP(hipMemsetAsync(b, 0, sizeof(double) * (size_t)nreal, st));
        ED_TRY(slab_from_rowmajor(*h, (const double *)d_f, own.p[0], st));
        F = own.p[0];
        if (want_vec) ev = own.p[1];
    }
    // evec = f keeps the start vector for the rebuild (HubMatrix_GPU.cuh:566)
    if (want_vec && !inplace)
        ED_HIP(hipMemcpyAsync(ev, F, sizeof(double) * (size_t)nreal, hipMemcpyDeviceToDevice, st));
    int64_t nb;
    ED_TRY(dot_partials(nreal, F, F, w.partial, &nb, st));
    ED_TRY(finalize_sum(w.partial, nb, w.bdev, 1, st));

    std::vector<double> a(maxN + 2, 0.0), araw(maxN + 2, 0.0), dl(maxN + 2, 0.0), b(maxN + 3, 0.0), z;
    double anorm = 0.0;
    int lastN = -1, status = ED_ENOCONV;
    int processed = 0; // a[0..processed-1] folded into anorm
    double *A = F, *B = w.buf[0], *C = w.buf[1];
    int N;
    bool stop = false;
    const bool vfree = slab && !reorth && !selective; // reorthogonalisation works on a stored v
    SelOrth so;
    so.n = nreal;
    so.cvec = cvec;
    for (N = 0; N < maxN && !stop; ++N) {
        if (vfree) {
            ED_TRY(lanczos_step_vfree(*h, nreal, A, B, C, w, N, st));
            double *t = B;
            B = A;
            A = C;
            C = t;
        } else {
            ED_TRY(lanczos_step(*h, cvec, nreal, A, B, C, w, N, reorth, slab, st));
            if (selective) ED_TRY(so.store(N, C, st)); // v_N
            double *t = A;
            A = B;
            B = C;
            C = t;
        }
        if (selective && !so.Y.empty()) {
            // f_{N+1} (in A) against the good Ritz vectors; b[N+1] recomputed
            int64_t nbk;
            ED_TRY(so.purge(A, w.partial, &nbk, st));
            ED_TRY(finalize_sum(w.partial, nbk, w.bdev + N + 1, 1, st));
            ++so.purges;
        }
        if (N % every == every - 1 || N == maxN - 1) {
            ED_HIP(hipMemcpyAsync(araw.data(), w.adev, sizeof(double) * (size_t)(N + 1), hipMemcpyDeviceToHost, st));
            ED_HIP(hipMemcpyAsync(b.data(), w.bdev, sizeof(double) * (size_t)(N + 2), hipMemcpyDeviceToHost, st));
            if (reorth)
                ED_HIP(hipMemcpyAsync(dl.data(), w.ddev, sizeof(double) * (size_t)(N + 1), hipMemcpyDeviceToHost, st));
            ED_HIP(hipStreamSynchronize(st));
            if (reorth)
                for (int k = 0; k <= N; ++k) a[k] = araw[k] + dl[k];
            else
                for (int k = 0; k <= N; ++k) a[k] = araw[k];
            if (selective && b[N + 1] > 0.0 && std::isfinite(a[N])) {
                if (nev > 1 && N + 1 >= nev) {
                    // keep only good Ritz vectors up to the nev-th lowest Ritz value (with a margin)
                    double tn = 0.0;
                    for (int k = 0; k <= N; ++k) tn = std::fmax(tn, std::fabs(a[k]) + b[k + 1] + (k ? b[k] : 0.0));
                    so.theta_cap = bisect_kth(N + 1, a.data(), b.data() + 1, nev - 1) +
                                   std::sqrt(std::numeric_limits<double>::epsilon()) * tn;
                }
                int found = 0;
                ED_TRY(so.analyse(N, a.data(), b.data(), st, &found));
                if (found) {
                    // v_N (in B) and f_{N+1} (in A) against every good Ritz vector; the kept copy
                    // of v_N follows
                    int64_t nbk;
                    ED_TRY(so.purge(B, w.partial, &nbk, st));
                    ED_TRY(so.purge(A, w.partial, &nbk, st));
                    ED_TRY(finalize_sum(w.partial, nbk, w.bdev + N + 1, 1, st));
                    ED_HIP(hipMemcpyAsync(so.vec(N), B, sizeof(double) * (size_t)nreal, hipMemcpyDeviceToDevice, st));
                    ED_HIP(hipMemcpyAsync(&b[N + 1], w.bdev + N + 1, sizeof(double), hipMemcpyDeviceToHost, st));
                    ED_HIP(hipStreamSynchronize(st));
                }
            }
            int M = N; // last usable step
            for (int k = processed; k <= N; ++k) {
                if (!(b[k + 1] > 0.0) || !std::isfinite(a[k])) { // breakdown / invariant subspace
                    M = k;
                    break;
                }
            }
            for (int k = processed; k <= M; ++k) {
                if (k == 0) anorm = std::fabs(a[0] + b[1]);
                else anorm = std::fmax(anorm, b[k] + std::fabs(a[k] + b[k + 1]));
            }
            processed = M + 1;
            z.assign(M + 1, 0.0);
            double lam;
            ed_tridiag_lowest(M + 1, a.data(), b.data() + 1, &lam, z.data());
            *eval = lam;
            lastN = M;
            // nev > 1: the nev lowest eigenvalues of T_{M+1} with copies of converged eigenvalues
            // merged and converging ghosts dropped (cw_lowest), each simple one passing the
            // residual test.  Selective orthogonalisation keeps most copies from forming, but not
            // around eigenvalues closer together than the accuracy at which a Ritz vector counts
            // as good (a pair 1e-8 apart in a disordered chain gave a third copy); it forms no
            // ghosts elsewhere, so a value the start vector barely touches is kept.
            bool rest = true;
            const bool exact = M < N || !(b[M + 1] > 0.0);
            const bool low_ok = std::fabs(b[M + 1] * z[M]) < tol * anorm && std::fabs(z[0]) > tol * anorm;
            // the filter (two tridiagonal eigenvalue solves) runs once the lowest value passes, and
            // at the last step for the values returned
            if (nev > 1 && (exact || low_ok || N == maxN - 1)) {
                std::vector<double> out((size_t)nev, std::numeric_limits<double>::quiet_NaN());
                const int f = cw_lowest(M + 1, a.data(), b.data() + 1, exact ? 0.0 : b[M + 1], nev, tol, anorm,
                                        true, out.data(), &rest);
                (void)f;
                for (int k = 0; k < nev; ++k) eval[k] = out[(size_t)k];
            }
            if (exact) {
                status = ED_OK; // exact invariant subspace: T_{M+1} eigenvalues are exact
                stop = true;
            } else if (rest && low_ok) {
                status = ED_OK;
                stop = true;
            }
        }
    }
    if (lastN < 0) return fail(ED_EINVAL, "ed_lanczos_eigen: no steps run");
    const int Nk = lastN;
    if (selective) {
        t_so_stats.good = (int)so.theta.size() / (cvec ? 2 : 1);
        t_so_stats.analyses = so.analyses;
        t_so_stats.purges = so.purges;
    } else if (Nk >= 1) {
        // the monitor's estimate of the loss of orthogonality over the run
        OmegaMonitor om;
        for (int j = 1; j <= Nk; ++j) t_so_stats.omega_max = std::fmax(t_so_stats.omega_max, om.step(a.data(), b.data()));
    }

    if (!want_vec) {
        ED_HIP(hipStreamSynchronize(st));
        if (a_out) memcpy(a_out, a.data(), sizeof(double) * (size_t)(Nk + 1));
        if (b_out) memcpy(b_out, b.data(), sizeof(double) * (size_t)(Nk + 2));
        if (nsteps) *nsteps = Nk + 1;
        if (status != ED_OK) return fail(status, "ed_lanczos_eigen: convergence test not met in %d steps", maxN);
        return ED_OK;
    }
    // Ritz vector: with the basis kept, V z; otherwise rerun the recurrence from the saved start
    // (HubMatrix_GPU.cuh:600-612)
    if (selective) {
        ED_TRY(so.combine(Nk, z.data(), ev, st));
    } else if (inplace) {
        A = w.buf[0];
        B = w.buf[1];
        C = (double *)d_evec;
        ED_TRY(slab_from_rowmajor(*h, (const double *)d_f, A, st));
        ev = (double *)d_f; // the start vector is in A now: f's buffer accumulates the Ritz vector
    } else {
        A = F;
        B = w.buf[0];
        C = w.buf[1];
        ED_HIP(hipMemcpyAsync(A, ev, sizeof(double) * (size_t)nreal, hipMemcpyDeviceToDevice, st));
    }
    if (!selective) ED_HIP(hipMemsetAsync(ev, 0, sizeof(double) * (size_t)nreal, st));
    for (int i = 0; !selective && i < Nk + 1; ++i) {
        if (vfree) { // update fused, no v
            ED_TRY(slab_lanczos_ritz_vfree(*h, A, B, C, w.bdev, w.adev, z[i], ev, i, st));
            double *t = B;
            B = A;
            A = C;
            C = t;
            continue;
        }
        if (slab) ED_TRY(slab_lanczos_ritz(*h, A, B, C, w.bdev, w.adev, z[i], ev, i, st)); // update fused
        else {
            ED_TRY(hv_lanczos_nodot(*h, cvec, A, B, C, B, w.bdev, i, st));
            ED_TRY(ritz_update(nreal, C, B, ev, w.adev, z[i], i, st));
        }
        if (reorth) {
            int64_t nbk;
            ED_TRY(lanczos_update(nreal, C, B, w.ddev, i, w.partial, &nbk, st));
        }
        double *t = A;
        A = B;
        B = C;
        C = t;
    }
    if (slab) {
        ED_TRY(slab_to_rowmajor(*h, ev, (double *)d_evec, st));
        ev = (double *)d_evec;
    }
    ED_TRY(ed_normalize(n, complex_vectors, ev, nullptr, stream));
    ED_HIP(hipStreamSynchronize(st));
    if (a_out) memcpy(a_out, a.data(), sizeof(double) * (size_t)(Nk + 1));
    if (b_out) memcpy(b_out, b.data(), sizeof(double) * (size_t)(Nk + 2));
    if (nsteps) *nsteps = Nk + 1;
    if (status != ED_OK) return fail(status, "ed_lanczos_eigen: convergence test not met in %d steps", maxN);
    return ED_OK;
}

extern "C" int ed_lanczos_eigen(ed_hubbard hh, int complex_vectors, const ed_lanczos_cfg *cfg, void *d_f,
                                double *eval, void *d_evec, double *a_out, double *b_out, int *nsteps, void *stream)
{
    return lanczos_eigen_impl(reinterpret_cast<const Hubbard *>(hh), complex_vectors, cfg, d_f, 1, eval, d_evec,
                              a_out, b_out, nsteps, stream);
}

extern "C" int ed_lanczos_lowest(ed_hubbard hh, int complex_vectors, const ed_lanczos_cfg *cfg, void *d_f, int nev,
                                 double *evals, void *d_evec, int *nsteps, void *stream)
{
    if (nev < 1 || nev > 64) return fail(ED_EINVAL, "ed_lanczos_lowest: nev must be 1..64");
    const Hubbard *h = reinterpret_cast<const Hubbard *>(hh);
    int n1 = 0;
    const int s = lanczos_eigen_impl(h, complex_vectors, cfg, d_f, nev, evals, d_evec, nullptr, nullptr, &n1, stream);
    if (nsteps) *nsteps = n1;
    if (s != ED_OK || !d_evec) return s;
    // The purges keep the basis semi-orthogonal, which leaves the eigenvalues accurate to working
    // precision but the Ritz vector V s_0 off by about sqrt(eps) along the purged directions.  A
    // plain Lanczos run (no reorthogonalisation) started from that vector, with the same test and
    // its Ritz vector rebuilt by recompute, removes that: its residual bounds the error.
    ed_lanczos_cfg c2{};
    if (cfg) c2 = *cfg;
    c2.reorth = 0;
    const bool cvec = complex_vectors != 0;
    const int64_t nreal = h->args.dim_up * h->args.dim_dn * (cvec ? 2 : 1);
    ED_HIP(hipMemcpyAsync(d_f, d_evec, sizeof(double) * (size_t)nreal, hipMemcpyDeviceToDevice, stream_of(stream)));
    double e0 = 0.0;
    int n2 = 0;
    const int s2 = lanczos_eigen_impl(h, complex_vectors, &c2, d_f, 1, &e0, d_evec, nullptr, nullptr, &n2, stream);
    if (nsteps) *nsteps = n1 + n2;
    if (s2 != ED_OK) return s2;
    // the refined vector is the Ritz vector of T's lowest eigenvalue: its value must be evals[0]
    // (a lowest value left out of evals by the spurious-value test would show up here)
    if (!(std::fabs(e0 - evals[0]) <= 1e-8 * std::fmax(1.0, std::fabs(e0))))
        return fail(ED_ENOCONV, "ed_lanczos_lowest: the ground state's value %.15g is not the lowest eigenvalue "
                                "found (%.15g)", e0, evals[0]);
    return ED_OK;
}

extern "C" int ed_lanczos_tridiagonal(ed_hubbard hh, int complex_vectors, int nsteps, const void *d_f,
                                      double *a_out, double *b_out, void *stream)
{
    const Hubbard *h = reinterpret_cast<const Hubbard *>(hh);
    if (!h || !d_f || nsteps < 1 || !a_out || !b_out) return fail(ED_EINVAL, "ed_lanczos_tridiagonal: bad arguments");
    const bool cvec = complex_vectors != 0;
    if (h->cplx_tables && !cvec) return fail(ED_EINVAL, "ed_lanczos_tridiagonal: complex tables need complex vectors");
    hipStream_t st = stream_of(stream);
    const int64_t n = h->args.dim_up * h->args.dim_dn;
    // real sectors run on the slab layout when the model has one (or gets one here, from 2^27
    // amplitudes on): only the coefficients leave, so f is converted once and nothing comes back
    if (!cvec) slab_auto(*const_cast<Hubbard *>(h), 3); // w.buf[0..1] and f
    const bool slab = h->slab && !cvec;
    const int64_t nreal = slab ? slab_len(*h) : (cvec ? 2 * n : n);
    Work w;
    ED_TRY(alloc_work(w, *h, nreal, nsteps, true));
    double *f = nullptr;
    ED_TRY(dmalloc((void **)&f, sizeof(double) * (size_t)nreal));
    int s = ED_OK;
    do {
        if (slab) {
            // padding entries stay zero in every vector of the recurrence
            for (double *b : {w.buf[0], w.buf[1], f})
                if (hipMemsetAsync(b, 0, sizeof(double) * (size_t)nreal, st) != hipSuccess) s = ED_EHIP;
            if (s) {
                s = fail(ED_EHIP, "ed_lanczos_tridiagonal: clear failed");
                break;
            }
            if ((s = slab_from_rowmajor(*h, (const double *)d_f, f, st))) break;
        } else if (hipMemcpyAsync(f, d_f, sizeof(double) * (size_t)nreal, hipMemcpyDeviceToDevice, st) !=
                   hipSuccess) {
            s = fail(ED_EHIP, "ed_lanczos_tridiagonal: copy failed");
            break;
        }
        int64_t nb;
        if ((s = dot_partials(nreal, f, f, w.partial, &nb, st))) break;
        if ((s = finalize_sum(w.partial, nb, w.bdev, 1, st))) break;
        double *A = f, *B = w.buf[0], *C = w.buf[1];
        for (int N = 0; N < nsteps && s == ED_OK; ++N) {
            if (slab) {
                s = lanczos_step_vfree(*h, nreal, A, B, C, w, N, st);
                double *t = B;
                B = A;
                A = C;
                C = t;
            } else {
                s = lanczos_step(*h, cvec, nreal, A, B, C, w, N, false, slab, st);
                double *t = A;
                A = B;
                B = C;
                C = t;
            }
        }
        if (s) break;
        if (hipMemcpyAsync(a_out, w.adev, sizeof(double) * (size_t)nsteps, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(b_out, w.bdev, sizeof(double) * (size_t)(nsteps + 1), hipMemcpyDeviceToHost, st) !=
                hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            s = fail(ED_EHIP, "ed_lanczos_tridiagonal: readback failed");
    } while (0);
    dfree(f);
    return s;
}

extern "C" int ed_continued_fraction(int n, const double *a, const double *b, double energy, double norm, int nw,
                                     const double *w_ri, double *g_ri)
{
    if (n < 1 || !a || !b || nw < 0 || (nw > 0 && (!w_ri || !g_ri)))
        return fail(ED_EINVAL, "ed_continued_fraction: bad arguments");
    for (int k = 0; k < nw; ++k) {
        const double wr = w_ri[2 * k] + energy, wi = w_ri[2 * k + 1];
        // g = 1/(w + E - a[n-1]) then g = 1/(w + E - a[i] - g b[i+1]^2) for i = n-2..0
        double gr, gi;
        {
            const double dr = wr - a[n - 1], di = wi;
            const double den = dr * dr + di * di;
            gr = dr / den;
            gi = -di / den;
        }
        for (int i = n - 2; i >= 0; --i) {
            const double bb = b[i + 1] * b[i + 1];
            const double dr = wr - a[i] - gr * bb, di = wi - gi * bb;
            const double den = dr * dr + di * di;
            gr = dr / den;
            gi = -di / den;
        }
        g_ri[2 * k] = norm * gr;
        g_ri[2 * k + 1] = norm * gi;
    }
    return ED_OK;
}

// ---- real-time evolution in the Krylov space (Krylov<C>, HubMatrix_GPU.cuh:619-728) ---------

namespace ed {

// One step of the propagation: phi += c q_i (complex c); then, unless this is the last term,
// q_{i+1} = ((y - b_i q_{i-1}) - a_i q_i) / b_{i+1} with y = H q_i, written over q_{i-1}.
__global__ __launch_bounds__(256) void k_krylov_step(int64_t n, double cr, double ci, double a, double b, double inv,
                                                     int last, const double *__restrict__ y, double *qprev,
                                                     const double *__restrict__ q, double *__restrict__ phi)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double qr = q[2 * i], qi = q[2 * i + 1];
    phi[2 * i] += cr * qr - ci * qi;
    phi[2 * i + 1] += cr * qi + ci * qr;
    if (last) return;
    const double tr = (y[2 * i] - b * qprev[2 * i]) - a * qr;
    const double ti = (y[2 * i + 1] - b * qprev[2 * i + 1]) - a * qi;
    qprev[2 * i] = tr * inv;
    qprev[2 * i + 1] = ti * inv;
}

static thread_local Arena t_krylov_arena;

} // namespace ed

extern "C" int ed_krylov_tridiagonal(ed_hubbard hh, const void *d_phi0, int max_steps, double btol, double *a_out,
                                     double *b_out, int *n_out, void *stream)
{
    const Hubbard *h = reinterpret_cast<const Hubbard *>(hh);
    if (!h || !d_phi0 || max_steps < 1 || !(btol >= 0.0) || !a_out || !b_out || !n_out)
        return fail(ED_EINVAL, "ed_krylov_tridiagonal: bad arguments");
    hipStream_t st = stream_of(stream);
    const int64_t n = h->args.dim_up * h->args.dim_dn, nreal = 2 * n;
    Work w;
    ED_TRY(alloc_work(w, *h, nreal, max_steps, true));
    double *f = nullptr;
    ED_TRY(dmalloc((void **)&f, sizeof(double) * (size_t)nreal));
    int s = ED_OK;
    std::vector<double> hb((size_t)max_steps + 1);
    int N = 0, nt = -1;
    do {
        // the input is used as it is (b_0 = 1: no normalisation), as Krylov::tridiagonal does
        const double one = 1.0;
        if (hipMemcpyAsync(f, d_phi0, sizeof(double) * (size_t)nreal, hipMemcpyDeviceToDevice, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess ||
            hipMemcpy(w.bdev, &one, sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
            s = fail(ED_EHIP, "ed_krylov_tridiagonal: setup failed");
            break;
        }
        double *A = f, *B = w.buf[0], *C = w.buf[1];
        while (N < max_steps && nt < 0 && s == ED_OK) {
            const int chunk = std::min(10, max_steps - N);
            for (int k = 0; k < chunk && s == ED_OK; ++k, ++N) {
                s = lanczos_step(*h, true, nreal, A, B, C, w, N, false, false, st);
                double *t = A;
                A = B;
                B = C;
                C = t;
            }
            if (s) break;
            // stop at the first b_k < btol (Krylov::tridiagonal breaks after appending it)
            if (hipStreamSynchronize(st) != hipSuccess ||
                hipMemcpy(hb.data(), w.bdev, sizeof(double) * (size_t)(N + 1), hipMemcpyDeviceToHost) != hipSuccess) {
                s = fail(ED_EHIP, "ed_krylov_tridiagonal: readback failed");
                break;
            }
            for (int k = N - chunk + 1; k <= N; ++k)
                if (hb[(size_t)k] < btol) {
                    nt = k;
                    break;
                }
        }
        if (s) break;
        if (nt < 0) nt = N;
        if (hipMemcpy(a_out, w.adev, sizeof(double) * (size_t)nt, hipMemcpyDeviceToHost) != hipSuccess) {
            s = fail(ED_EHIP, "ed_krylov_tridiagonal: readback failed");
            break;
        }
        b_out[0] = 0.0;
        for (int k = 1; k <= nt; ++k) b_out[k] = hb[(size_t)k];
        *n_out = nt;
    } while (0);
    (void)hipStreamSynchronize(st);
    dfree(f);
    return s;
}

extern "C" int ed_krylov_time_evol(ed_hubbard hh, const void *d_phi0, int n, const double *a, const double *b,
                                   const double *m, const double *z, double dt, void *d_phi, void *stream)
{
    const Hubbard *h = reinterpret_cast<const Hubbard *>(hh);
    if (!h || !d_phi0 || !d_phi || n < 1 || !a || !b || !m || !z)
        return fail(ED_EINVAL, "ed_krylov_time_evol: bad arguments");
    for (int i = 1; i < n; ++i)
        if (!(b[i] != 0.0)) return fail(ED_EINVAL, "ed_krylov_time_evol: b[%d] is zero", i);
    hipStream_t st = stream_of(stream);
    const int64_t D = h->args.dim_up * h->args.dim_dn;
    const size_t bytes = sizeof(double) * 2 * (size_t)D;
    // coefficients of the Krylov vectors: c_i = sum_j exp(-i m_j dt) z_0j z_ij
    std::vector<double> cr((size_t)n, 0.0), ci((size_t)n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            const double zz = z[(size_t)j * n] * z[(size_t)j * n + i];
            cr[(size_t)i] += std::cos(m[j] * dt) * zz;
            ci[(size_t)i] += -std::sin(m[j] * dt) * zz;
        }
    Scratch S(t_krylov_arena, st);
    ED_TRY(S.reserve(3 * Scratch::round(bytes)));
    double *q = (double *)S.take(bytes), *qp = (double *)S.take(bytes), *y = (double *)S.take(bytes);
    if (hipMemcpyAsync(q, d_phi0, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess ||
        hipMemsetAsync(qp, 0, bytes, st) != hipSuccess)
        return fail(ED_EHIP, "ed_krylov_time_evol: setup failed");
    const unsigned grid = (unsigned)((D + 255) / 256);
    for (int i = 0; i < n; ++i) {
        const bool last = i + 1 == n;
        if (!last) ED_TRY(hv_plain(*h, true, q, y, st));
        hipLaunchKernelGGL(k_krylov_step, dim3(grid), dim3(256), 0, st, D, cr[(size_t)i], ci[(size_t)i], a[i], b[i],
                           last ? 0.0 : 1.0 / b[i + 1], last ? 1 : 0, (const double *)y, qp, (const double *)q,
                           (double *)d_phi);
        ED_CHECK_LAUNCH();
        std::swap(q, qp);
    }
    return ED_OK;
}

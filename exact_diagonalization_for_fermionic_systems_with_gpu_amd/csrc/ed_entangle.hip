// ed_entangle.hip -- entanglement entropy of psi from particle-number-block reduced
// density matrices.
//
// Reference: EEmatHubAB_GPU::Make_Write + make_write_ (MBL_test/include/
// entangle_entropy_GPU.cuh:27-144) scatter psi into one dense rowD x colD matrix indexed by
// (A_up, A_dn) x (B_up, B_dn) sub-states; cusolverDsvd/Zsvd (entangle_entropy.h:231-353)
// takes its singular values and the drivers sum S = -sum_{s != 0} s ln s
// (Hubbard_EE_test.cu:78-83).  That matrix is block-sparse: psi[u,d] only connects rows
// and columns with the same particle numbers (n_A^up, n_A^dn) in A, so its singular values
// are the union of the blocks' singular values, whatever order the reference numbers its
// sub-states in.  Here:
//   1. one gather pass over psi writes every block contiguously, with the smaller block
//      dimension as rows (so the product below is the small side);
//   2. rho = Psi Psi^T (Psi Psi^H) per block with a hand-written FP64 MFMA kernel
//      (v_mfma_f64_16x16x4f64, 64x64 output tiles, upper-triangle tiles only);
//   3. rocSOLVER dsyevd/zheevd (values only) gives lambda -> S_vN = -sum lambda ln lambda;
//   4. rocSOLVER dgesvd/zgesvd (values only) on the block itself gives the Schmidt values
//      sigma -> S_ref = -sum sigma ln sigma, the reference's formula.  sigma is taken from
//      the SVD rather than sqrt(lambda) because the entropy's small-sigma terms need sigma
//      to absolute accuracy eps*|Psi|, which the eigenvalues of rho (accurate to eps*|rho|)
//      do not give.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ed_common.hpp"
#include "ed_thread.hpp"

namespace ed {

// real rho_A blocks of this many rows and more take the two-stage eigenvalue routine
// (ed_symeig.hip) instead of rocSOLVER's one-stage dsyevd
constexpr int64_t kTwoStageMin = 1024;
int sym_eigvals_two_stage(rocblas_handle h, hipStream_t st, int64_t n, double *A, int64_t lda, bool upper_only,
                          double *w);
struct SymGroup { // ed_symeig.hip
    int64_t n, lda;
    int count;
    double *A, *w;
};
int sym_eigvals_groups(rocblas_handle h, hipStream_t st, const std::vector<SymGroup> &groups, bool upper_only);

typedef double dbl4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------------
// 1. gather psi into the blocks
// ---------------------------------------------------------------------------------
struct BlockGeo {
    int64_t off;   // element offset of the block in the buffer
    int64_t rows;  // stored rows (the smaller dimension)
    int64_t cols;
    int32_t ma_dn; // C(|A|, nA_dn): A index = ra_up*ma_dn + ra_dn
    int32_t nb_dn; // C(|B|, Q_dn - nA_dn): B index = rb_up*nb_dn + rb_dn
    int32_t trans; // 1: stored as (B index, A index)
    int32_t pad;
};

template <bool CPLX>
__global__ __launch_bounds__(256) void k_gather_blocks(int64_t dim_up, int64_t dim_dn,
                                                       const int32_t *__restrict__ cu,  // per u: blk, ra, rb
                                                       const int32_t *__restrict__ cd,  // per d: blk, ra, rb
                                                       int nbd, const BlockGeo *__restrict__ geo,
                                                       const double *__restrict__ psi, double *__restrict__ out)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= dim_up * dim_dn) return;
    const int64_t u = g / dim_dn, d = g - u * dim_dn;
    const int bu = cu[3 * u], rau = cu[3 * u + 1], rbu = cu[3 * u + 2];
    const int bd = cd[3 * d], rad = cd[3 * d + 1], rbd = cd[3 * d + 2];
    const BlockGeo G = geo[bu * nbd + bd];
    const int64_t ia = (int64_t)rau * G.ma_dn + rad, ib = (int64_t)rbu * G.nb_dn + rbd;
    const int64_t idx = G.off + (G.trans ? ib * G.cols + ia : ia * G.cols + ib);
    if (CPLX) {
        out[2 * idx] = psi[2 * g];
        out[2 * idx + 1] = psi[2 * g + 1];
    } else {
        out[idx] = psi[g];
    }
}

// ---------------------------------------------------------------------------------
// 2. rho = A A^T (A A^H), A row-major k x K (lda = K), rho column-major k x k (ld = k),
//    upper-triangle tiles only.  256 threads = 4 waves, 64x64 tile per workgroup, each
//    wave a 32x32 quadrant = 2x2 MFMA 16x16 tiles; K staged through two LDS stages 16 at a
//    time, the next 16 loaded into registers while the MFMAs of the current 16 run.  (128x128 tiles,
//    at one wave per SIMD, measured no faster at 18 sites: 0.623 against 0.618 s.)
//    MFMA f64 16x16x4 operand layout: A lane l -> (row l%16, k l/16); B lane l ->
//    (col l%16, k l/16); D lane l -> (rows l/16 + 4r, col l%16), r = 0..3 (checked on
//    gfx950 with a one-wave decode of (i+1)(j+1) products).
// ---------------------------------------------------------------------------------
constexpr int TT = 64, TK = 16;

template <bool CPLX>
__global__ __launch_bounds__(256) void k_gram_mfma(int64_t k, int64_t K, const double *__restrict__ A,
                                                   double *__restrict__ rho, int ntile, int64_t sA, int64_t sR)
{
    // blockIdx.y: the state of a batch (block matrices sA, rho matrices sR doubles apart)
    A += (int64_t)blockIdx.y * sA;
    rho += (int64_t)blockIdx.y * sR;
    // map blockIdx -> (ti <= tj)
    int64_t b = blockIdx.x, ti = 0;
    {
        int64_t rem = b, row_len = ntile;
        while (rem >= row_len) {
            rem -= row_len;
            ++ti;
            --row_len;
        }
        b = ti + rem; // tj
    }
    const int64_t tj = b;
    const int64_t i0 = ti * TT, j0 = tj * TT;
    constexpr int NC = CPLX ? 2 : 1;
    __shared__ double sa[2][NC][TT][TK + 1]; // two stages: one is read by the MFMAs while the
    __shared__ double sb[2][NC][TT][TK + 1]; // next chunk is stored into the other
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int NP = TT / 32;
    const int wi = (wave >> 1) * (TT / 2), wj = (wave & 1) * (TT / 2);
    dbl4 acc[NP][NP], acci[NP][NP];
#pragma unroll
    for (int p = 0; p < NP; ++p)
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            acc[p][q] = dbl4{0, 0, 0, 0};
            acci[p][q] = dbl4{0, 0, 0, 0};
        }
    // the next 64 x 16 chunks of rows i0.. and j0.. are loaded into registers while the MFMAs
    // of the current chunk run (4 elements per thread per matrix; zero outside the block)
    constexpr int EPT = TT * TK / 256;
    double ra[EPT][NC], rb[EPT][NC];
    auto fetch = [&](int64_t k0) {
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
            const int e = tid + q * 256, r = e / TK, c = e % TK;
            const int64_t kk = k0 + c, ri = i0 + r, rj = j0 + r;
#pragma unroll
            for (int z = 0; z < NC; ++z) {
                ra[q][z] = (ri < k && kk < K) ? A[(ri * K + kk) * NC + z] : 0.0;
                rb[q][z] = (rj < k && kk < K) ? A[(rj * K + kk) * NC + z] : 0.0;
            }
        }
    };
    auto stash = [&](int st) {
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
            const int e = tid + q * 256, r = e / TK, c = e % TK;
#pragma unroll
            for (int z = 0; z < NC; ++z) {
                sa[st][z][r][c] = ra[q][z];
                sb[st][z][r][c] = rb[q][z];
            }
        }
    };
    fetch(0);
    stash(0);
    __syncthreads();
    int cur = 0;
    for (int64_t k0 = 0; k0 < K; k0 += TK) {
        const bool more = k0 + TK < K;
        if (more) fetch(k0 + TK);
#pragma unroll
        for (int kk = 0; kk < TK; kk += 4) {
            const int kl = kk + (lane >> 4), rl = lane & 15;
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                const double ar = sa[cur][0][wi + p * 16 + rl][kl];
                const double ai = CPLX ? sa[cur][NC - 1][wi + p * 16 + rl][kl] : 0.0;
#pragma unroll
                for (int q = 0; q < NP; ++q) {
                    const double br = sb[cur][0][wj + q * 16 + rl][kl];
                    acc[p][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, acc[p][q], 0, 0, 0);
                    if (CPLX) {
                        const double bi = sb[cur][NC - 1][wj + q * 16 + rl][kl];
                        acc[p][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bi, acc[p][q], 0, 0, 0);
                        acci[p][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, br, acci[p][q], 0, 0, 0);
                        acci[p][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ar, bi, acci[p][q], 0, 0, 0);
                    }
                }
            }
        }
        // one barrier per chunk: the stage just read is overwritten only after the next barrier
        if (more) stash(cur ^ 1);
        __syncthreads();
        cur ^= 1;
    }
    // store: rho(i, j) column-major at i + j*k
#pragma unroll
    for (int p = 0; p < NP; ++p)
#pragma unroll
        for (int q = 0; q < NP; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t i = i0 + wi + p * 16 + (lane >> 4) + 4 * r;
                const int64_t j = j0 + wj + q * 16 + (lane & 15);
                if (i < k && j < k) {
                    if (CPLX) {
                        rho[2 * (i + j * k)] = acc[p][q][r];
                        rho[2 * (i + j * k) + 1] = acci[p][q][r];
                    } else {
                        rho[i + j * k] = acc[p][q][r];
                    }
                }
            }
}

// ---------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------
static uint32_t pext32(uint32_t s, uint32_t mask)
{
    uint32_t r = 0;
    int k = 0;
    for (int i = 0; i < 32; ++i)
        if ((mask >> i) & 1u) {
            if ((s >> i) & 1u) r |= 1u << k;
            ++k;
        }
    return r;
}

static int64_t colex_rank(uint32_t s)
{
    int64_t r = 0;
    int k = 0;
    for (int p = 0; p < 32; ++p)
        if ((s >> p) & 1u) r += binom(p, ++k);
    return r;
}

struct Plan {
    int nA, nB;
    int amin_u, amax_u, amin_d, amax_d, nbu, nbd;
    std::vector<BlockGeo> geo;
    std::vector<int32_t> cu, cd; // per sector state: block index, A rank, B rank
    int64_t total = 0;
};

static void species_codes(int N, int Q, uint32_t maskA, int amin, std::vector<int32_t> &code)
{
    const uint32_t full = (N == 32) ? 0xffffffffu : ((1u << N) - 1u), maskB = full & ~maskA;
    code.clear();
    for (uint32_t s = 0; s <= full; ++s) {
        if (__builtin_popcount(s) != Q) {
            if (s == full) break;
            continue;
        }
        const uint32_t a = pext32(s, maskA), b = pext32(s, maskB);
        code.push_back(__builtin_popcount(a) - amin);
        code.push_back((int32_t)colex_rank(a));
        code.push_back((int32_t)colex_rank(b));
        if (s == full) break;
    }
}

static int make_plan(int N, int Qu, int Qd, uint32_t maskA, Plan &P)
{
    P.nA = __builtin_popcount(maskA);
    P.nB = N - P.nA;
    P.amin_u = std::max(0, Qu - P.nB);
    P.amax_u = std::min(Qu, P.nA);
    P.amin_d = std::max(0, Qd - P.nB);
    P.amax_d = std::min(Qd, P.nA);
    P.nbu = P.amax_u - P.amin_u + 1;
    P.nbd = P.amax_d - P.amin_d + 1;
    P.geo.assign((size_t)P.nbu * P.nbd, BlockGeo{});
    int64_t off = 0;
    for (int bu = 0; bu < P.nbu; ++bu)
        for (int bd = 0; bd < P.nbd; ++bd) {
            const int au = P.amin_u + bu, ad = P.amin_d + bd;
            const int64_t ma = binom(P.nA, au) * binom(P.nA, ad);
            const int64_t nb = binom(P.nB, Qu - au) * binom(P.nB, Qd - ad);
            BlockGeo &G = P.geo[(size_t)bu * P.nbd + bd];
            G.off = off;
            G.trans = ma > nb ? 1 : 0;
            G.rows = std::min(ma, nb);
            G.cols = std::max(ma, nb);
            G.ma_dn = (int32_t)binom(P.nA, ad);
            G.nb_dn = (int32_t)binom(P.nB, Qd - ad);
            off += ma * nb;
        }
    P.total = off;
    species_codes(N, Qu, maskA, P.amin_u, P.cu);
    species_codes(N, Qd, maskA, P.amin_d, P.cd);
    return ED_OK;
}

// the calling thread's rocBLAS handle (ed_thread.hpp: one per host thread, destroyed at exit)
static rocblas_handle solver_handle() { return thread_blas_handle(); }

// worker streams (default priority) and rocSOLVER handles of one device for the concurrent block
// solves, created on first use and kept until the process ends; calls that use them take turns
constexpr int kEeWorkers = 4;
struct EePool {
    hipStream_t st[kEeWorkers] = {};
    rocblas_handle h[kEeWorkers] = {};
    bool warm[2][2] = {}; // [complex][with Schmidt values]: routines already run on a calling thread
};

// a k x k symmetric (Hermitian) test matrix, column-major: 1/(1+|i-j|) (+ i 0.1 sign(j-i))
__global__ void k_warm_matrix(int64_t k, int cplx, double *A)
{
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= k * k) return;
    const int64_t i = q % k, j = q / k;
    const double re = 1.0 / (1.0 + (double)(i > j ? i - j : j - i));
    if (cplx) {
        A[2 * q] = re;
        A[2 * q + 1] = i < j ? 0.1 : (i > j ? -0.1 : 0.0);
    } else {
        A[q] = re;
    }
}
static std::mutex g_ee_mu;
static EePool *g_ee_pool[64] = {};

// the pool of the current device (g_ee_mu held); nullptr if it cannot be made
static EePool *ee_pool(int dev)
{
    if (dev < 0 || dev >= 64) return nullptr;
    if (g_ee_pool[dev]) return g_ee_pool[dev];
    EePool *p = new EePool;
    bool ok = true;
    for (int i = 0; i < kEeWorkers && ok; ++i) {
        ok = hipStreamCreateWithFlags(&p->st[i], hipStreamNonBlocking) == hipSuccess &&
             rocblas_create_handle(&p->h[i]) == rocblas_status_success &&
             rocblas_set_stream(p->h[i], p->st[i]) == rocblas_status_success;
    }
    if (!ok) {
        for (int i = 0; i < kEeWorkers; ++i) {
            if (p->h[i]) rocblas_destroy_handle(p->h[i]);
            if (p->st[i]) (void)hipStreamDestroy(p->st[i]);
        }
        delete p;
        (void)hipGetLastError();
        return nullptr;
    }
    g_ee_pool[dev] = p;
    return p;
}

// kernel-time statistics of the block spectra, summed over calls until reset: the Gram kernels'
// durations and their SYRK FLOPs, and the durations of the eigenvalue (and SVD) stage, each
// measured by events on the stream that ran it
struct EeStats {
    double gram_ms = 0.0, gram_flops = 0.0, eig_ms = 0.0;
};
static std::mutex g_stats_mu;
static EeStats g_stats;

// spectra of all blocks: lambda (rho eigenvalues, clipped at 0) and sigma (Schmidt values)
static int block_spectra(int N, int Qu, int Qd, int64_t maskA, bool cplx, const void *d_psi, hipStream_t st,
                         std::vector<double> &lambda, std::vector<double> &sigma, bool want_sigma)
{
    if (N < 2 || N > 31 || Qu < 0 || Qd < 0 || Qu > N || Qd > N || !d_psi)
        return fail(ED_EINVAL, "entanglement: bad arguments");
    const uint32_t full = (1u << N) - 1u;
    const uint32_t mA = (uint32_t)maskA & full;
    if (mA == 0 || mA == full) return fail(ED_EINVAL, "entanglement: A must be a proper subset of the sites");
    Plan P;
    ED_TRY(make_plan(N, Qu, Qd, mA, P));
    const int64_t dim_up = (int64_t)P.cu.size() / 3, dim_dn = (int64_t)P.cd.size() / 3;
    const int NC = cplx ? 2 : 1;
    rocblas_handle h = solver_handle();
    if (!h) return fail(ED_EHIP, "entanglement: rocblas handle creation failed");
    rocblas_set_stream(h, st);

    double *dM = nullptr, *dR = nullptr, *dWall = nullptr, *dSall = nullptr, *dE = nullptr;
    int32_t *dcu = nullptr, *dcd = nullptr;
    BlockGeo *dgeo = nullptr;
    rocblas_int *dinfo = nullptr;
    int64_t maxk = 1, maxK = 1, nvals = 0;
    std::vector<int> jobs;       // nonempty blocks, plan order
    std::vector<int64_t> woff;   // offset of each job's values in dWall / dSall
    for (int g = 0; g < (int)P.geo.size(); ++g) {
        const BlockGeo &G = P.geo[g];
        if (G.rows == 0 || G.cols == 0) continue;
        maxk = std::max(maxk, G.rows);
        maxK = std::max(maxK, G.cols);
        jobs.push_back(g);
        woff.push_back(nvals);
        nvals += G.rows;
    }
    const int njobs = (int)jobs.size();
    std::vector<int> order(njobs);
    for (int b = 0; b < njobs; ++b) order[b] = b;
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return P.geo[jobs[x]].rows > P.geo[jobs[y]].rows; });
    // when the largest block has 2048 rows or more (ED_EE_PARALLEL_MIN), the blocks are solved
    // kEeWorkers at a time (rocSOLVER's tridiagonal reduction is launch- and matrix-vector bound:
    // one alone leaves most of the GPU idle).  Each block's arithmetic is the same either way.
    const char *pm = getenv("ED_EE_PARALLEL_MIN");
    const int64_t par_min = pm ? std::max<int64_t>(1, atoll(pm)) : 2048;
    int dev = 0;
    ED_HIP(hipGetDevice(&dev));
    EePool *pool = nullptr;
    std::unique_lock<std::mutex> pool_lock;
    int nw = 1;
    if (maxk >= par_min && njobs > 1) {
        pool_lock = std::unique_lock<std::mutex>(g_ee_mu);
        pool = ee_pool(dev);
        if (pool) nw = std::min(kEeWorkers, njobs);
        else pool_lock.unlock(); // no worker streams: one block at a time on st
        // each worker holds one rho block: fewer workers when their blocks would not fit
        size_t free_b = 0, total_b = 0;
        if (nw > 1 && hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
            const double blk = (double)(sizeof(double) * NC) * (double)maxk * (double)maxk;
            const double rest = (double)(sizeof(double) * NC) * (double)P.total;
            while (nw > 1 && rest + nw * blk > 0.8 * (double)free_b) --nw;
        }
    }
    // a worker's failure is reported by the calling thread (the error text is per thread)
    std::vector<std::string> wmsg((size_t)nw);
    auto wfail = [&](int wi, const char *msg, int code) {
        char buf[160];
        snprintf(buf, sizeof buf, msg, code);
        wmsg[wi] = buf;
    };
    hipEvent_t ready = nullptr;
    ED_HIP(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    struct EvGuard {
        hipEvent_t e;
        ~EvGuard() { (void)hipEventDestroy(e); }
    } ev_guard{ready};
    int s = ED_OK;
    static thread_local Arena t_ee_arena;
    Scratch S(t_ee_arena, st);
    do {
        const size_t bM = sizeof(double) * NC * (size_t)P.total,
                     bR = Scratch::round(sizeof(double) * NC * (size_t)(maxk * maxk)),
                     bV = sizeof(double) * (size_t)std::max<int64_t>(nvals, 1),
                     bE = Scratch::round(sizeof(double) * (size_t)std::max(maxk, maxK)),
                     bU = sizeof(int32_t) * P.cu.size(), bD = sizeof(int32_t) * P.cd.size(),
                     bG = sizeof(BlockGeo) * P.geo.size(),
                     bI = sizeof(rocblas_int) * (size_t)std::max(2 * njobs, 1);
        if ((s = S.reserve(Scratch::round(bM) + nw * bR + 2 * Scratch::round(bV) + nw * bE + Scratch::round(bU) +
                           Scratch::round(bD) + Scratch::round(bG) + Scratch::round(bI))))
            break;
        dM = (double *)S.take(bM);
        dR = (double *)S.take(nw * bR);
        dWall = (double *)S.take(bV);
        dSall = (double *)S.take(bV);
        dE = (double *)S.take(nw * bE);
        dcu = (int32_t *)S.take(bU);
        dcd = (int32_t *)S.take(bD);
        dgeo = (BlockGeo *)S.take(bG);
        dinfo = (rocblas_int *)S.take(bI);
        if (hipMemcpyAsync(dcu, P.cu.data(), sizeof(int32_t) * P.cu.size(), hipMemcpyHostToDevice, st) ||
            hipMemcpyAsync(dcd, P.cd.data(), sizeof(int32_t) * P.cd.size(), hipMemcpyHostToDevice, st) ||
            hipMemcpyAsync(dgeo, P.geo.data(), sizeof(BlockGeo) * P.geo.size(), hipMemcpyHostToDevice, st)) {
            s = fail(ED_EHIP, "entanglement: upload failed");
            break;
        }
        const int64_t D = dim_up * dim_dn;
        const unsigned grid = (unsigned)((D + 255) / 256);
        if (cplx)
            hipLaunchKernelGGL(k_gather_blocks<true>, dim3(grid), dim3(256), 0, st, dim_up, dim_dn, dcu, dcd, P.nbd,
                               dgeo, (const double *)d_psi, dM);
        else
            hipLaunchKernelGGL(k_gather_blocks<false>, dim3(grid), dim3(256), 0, st, dim_up, dim_dn, dcu, dcd,
                               P.nbd, dgeo, (const double *)d_psi, dM);
        if (hipGetLastError() != hipSuccess) {
            s = fail(ED_EHIP, "entanglement: gather launch failed");
            break;
        }
        if (hipEventRecord(ready, st) != hipSuccess) {
            s = fail(ED_EHIP, "entanglement: event record failed");
            break;
        }
        // Real blocks of kTwoStageMin rows and more (S_vN only): every one's rho resident at once,
        // the blocks of one size reduced together (strided-batched panels and updates) and the
        // bulge chasing and bisection of all of them in shared launches (ed_symeig.hip), on st.
        // The other blocks go to the workers below.
        if (!cplx && !want_sigma) {
            std::map<int64_t, std::vector<int>> big; // rows -> jobs
            double rho_doubles = 0.0, w_doubles = 0.0;
            for (int b = 0; b < njobs; ++b) {
                const int64_t k = P.geo[jobs[b]].rows;
                if (k < kTwoStageMin) continue;
                big[k].push_back(b);
                rho_doubles += (double)k * (double)k;
                w_doubles += (double)k;
            }
            size_t free_b = 0, total_b = 0;
            if (!big.empty() && hipMemGetInfo(&free_b, &total_b) == hipSuccess &&
                8.0 * (rho_doubles + w_doubles) < 0.7 * (double)free_b) {
                static thread_local Arena t_eeg_arena;
                Scratch SG(t_eeg_arena, st);
                if ((s = SG.reserve(Scratch::round(sizeof(double) * (size_t)rho_doubles) +
                                    Scratch::round(sizeof(double) * (size_t)w_doubles))))
                    break;
                double *rho_all = (double *)SG.take(sizeof(double) * (size_t)rho_doubles);
                double *w_all = (double *)SG.take(sizeof(double) * (size_t)w_doubles);
                std::vector<SymGroup> groups;
                std::vector<std::pair<int, double *>> wslot; // job -> its eigenvalues in w_all
                hipEvent_t te[3] = {nullptr, nullptr, nullptr};
                for (hipEvent_t &e : te)
                    if (hipEventCreate(&e) != hipSuccess) s = fail(ED_EHIP, "entanglement: event creation failed");
                struct TeGuard {
                    hipEvent_t *e;
                    ~TeGuard()
                    {
                        for (int i = 0; i < 3; ++i)
                            if (e[i]) (void)hipEventDestroy(e[i]);
                    }
                } te_guard{te};
                if (s) break;
                double flops = 0.0;
                (void)hipEventRecord(te[0], st);
                double *rp = rho_all, *wp = w_all;
                for (auto &kv : big) {
                    const int64_t k = kv.first;
                    SymGroup g{k, k, (int)kv.second.size(), rp, wp};
                    for (int b : kv.second) {
                        const BlockGeo &G = P.geo[jobs[b]];
                        const int ntile = (int)((k + TT - 1) / TT);
                        const int64_t nblk = (int64_t)ntile * (ntile + 1) / 2;
                        hipLaunchKernelGGL(k_gram_mfma<false>, dim3((unsigned)nblk), dim3(256), 0, st, k, G.cols,
                                           (const double *)(dM + G.off), rp, ntile, (int64_t)0, (int64_t)0);
                        flops += (double)k * (double)(k + 1) * (double)G.cols;
                        wslot.push_back({b, wp});
                        rp += k * k;
                        wp += k;
                    }
                    groups.push_back(g);
                }
                if (hipGetLastError() != hipSuccess) {
                    s = fail(ED_EHIP, "entanglement: gram launch failed");
                    break;
                }
                (void)hipEventRecord(te[1], st);
                if ((s = sym_eigvals_groups(h, st, groups, true))) break;
                (void)hipEventRecord(te[2], st);
                for (auto &bw : wslot) {
                    const int64_t k = P.geo[jobs[bw.first]].rows;
                    if (hipMemcpyAsync(dWall + woff[bw.first], bw.second, sizeof(double) * (size_t)k,
                                       hipMemcpyDeviceToDevice, st) != hipSuccess ||
                        hipMemsetAsync(dinfo + bw.first, 0, sizeof(rocblas_int), st) != hipSuccess) {
                        s = fail(ED_EHIP, "entanglement: eigenvalue copy failed");
                        break;
                    }
                }
                if (s) break;
                if (hipEventSynchronize(te[2]) != hipSuccess) {
                    s = fail(ED_EHIP, "entanglement: two-stage eigenvalues failed");
                    break;
                }
                float gms = 0.0f, ems = 0.0f;
                (void)hipEventElapsedTime(&gms, te[0], te[1]);
                (void)hipEventElapsedTime(&ems, te[1], te[2]);
                {
                    std::lock_guard<std::mutex> lk(g_stats_mu);
                    g_stats.gram_ms += gms;
                    g_stats.gram_flops += flops;
                    g_stats.eig_ms += ems;
                }
                std::vector<int> rest;
                for (int b : order)
                    if (P.geo[jobs[b]].rows < kTwoStageMin) rest.push_back(b);
                order.swap(rest);
            }
        }
        // the blocks go to the workers largest first; each worker runs Gram, syevd (and gesvd)
        // on its own stream, rho buffer and rocSOLVER handle, and leaves the values in its
        // block's slot of dWall / dSall
        std::atomic<int> next(0);
        const int nwork = (int)order.size();
        auto worker = [&](int wi, hipStream_t ws, rocblas_handle hw) {
            if (nw > 1 && hipSetDevice(dev) != hipSuccess) {
                wfail(wi, "entanglement: worker device selection failed (%d)", 0);
                return;
            }
            if (hipStreamWaitEvent(ws, ready, 0) != hipSuccess) {
                wfail(wi, "entanglement: worker stream wait failed (%d)", 0);
                return;
            }
            double *rho = dR + (size_t)wi * (bR / sizeof(double)), *E = dE + (size_t)wi * (bE / sizeof(double));
            // timing events of this worker's blocks: before the Gram, after it, after the solvers
            std::vector<hipEvent_t> tev;
            double flops = 0.0;
            struct TevGuard {
                std::vector<hipEvent_t> &v;
                ~TevGuard()
                {
                    for (hipEvent_t e : v) (void)hipEventDestroy(e);
                }
            } tev_guard{tev};
            auto mark = [&]() {
                hipEvent_t e = nullptr;
                if (hipEventCreate(&e) != hipSuccess) return false;
                tev.push_back(e);
                return hipEventRecord(e, ws) == hipSuccess;
            };
            for (int q = next.fetch_add(1); q < nwork; q = next.fetch_add(1)) {
                const int b = order[q];
                const BlockGeo &G = P.geo[jobs[b]];
                const int64_t k = G.rows, K = G.cols;
                double *A = dM + NC * G.off;
                const int ntile = (int)((k + TT - 1) / TT);
                const int64_t nblk = (int64_t)ntile * (ntile + 1) / 2;
                if (!mark()) {
                    wfail(wi, "entanglement: timing event failed (%d)", 0);
                    return;
                }
                flops += (cplx ? 4.0 : 1.0) * (double)k * (double)(k + 1) * (double)K;
                if (cplx)
                    hipLaunchKernelGGL(k_gram_mfma<true>, dim3((unsigned)nblk), dim3(256), 0, ws, k, K, A, rho, ntile,
                                       (int64_t)0, (int64_t)0);
                else
                    hipLaunchKernelGGL(k_gram_mfma<false>, dim3((unsigned)nblk), dim3(256), 0, ws, k, K, A, rho, ntile,
                                       (int64_t)0, (int64_t)0);
                if (hipGetLastError() != hipSuccess || !mark()) {
                    wfail(wi, "entanglement: gram launch failed (%d)", 0);
                    return;
                }
                rocblas_status rs = rocblas_status_success;
                if (cplx)
                    rs = rocsolver_zheevd(hw, rocblas_evect_none, rocblas_fill_upper, (rocblas_int)k,
                                          (rocblas_double_complex *)rho, (rocblas_int)k, dWall + woff[b], E,
                                          dinfo + b);
                else if (k >= kTwoStageMin) {
                    // large real blocks: two-stage reduction (ed_symeig.hip); its info slot stays 0
                    if (hipMemsetAsync(dinfo + b, 0, sizeof(rocblas_int), ws) != hipSuccess ||
                        sym_eigvals_two_stage(hw, ws, k, rho, k, true, dWall + woff[b]) != ED_OK) {
                        wfail(wi, "entanglement: two-stage eigenvalues failed (%d)", 0);
                        return;
                    }
                } else
                    rs = rocsolver_dsyevd(hw, rocblas_evect_none, rocblas_fill_upper, (rocblas_int)k, rho,
                                          (rocblas_int)k, dWall + woff[b], E, dinfo + b);
                if (rs != rocblas_status_success) {
                    wfail(wi, "entanglement: rocsolver syevd failed (%d)", (int)rs);
                    return;
                }
                if (!want_sigma) {
                    if (!mark()) {
                        wfail(wi, "entanglement: timing event failed (%d)", 0);
                        return;
                    }
                    continue;
                }
                // Schmidt values: SVD of the block (column-major K x k, ld K); A is overwritten
                if (cplx)
                    rs = rocsolver_zgesvd(hw, rocblas_svect_none, rocblas_svect_none, (rocblas_int)K, (rocblas_int)k,
                                          (rocblas_double_complex *)A, (rocblas_int)K, dSall + woff[b], nullptr, 1,
                                          nullptr, 1, E, rocblas_outofplace, dinfo + njobs + b);
                else
                    rs = rocsolver_dgesvd(hw, rocblas_svect_none, rocblas_svect_none, (rocblas_int)K, (rocblas_int)k,
                                          A, (rocblas_int)K, dSall + woff[b], nullptr, 1, nullptr, 1, E,
                                          rocblas_outofplace, dinfo + njobs + b);
                if (rs != rocblas_status_success) {
                    wfail(wi, "entanglement: rocsolver gesvd failed (%d)", (int)rs);
                    return;
                }
                if (!mark()) {
                    wfail(wi, "entanglement: timing event failed (%d)", 0);
                    return;
                }
            }
            if (tev.empty()) return;
            if (hipEventSynchronize(tev.back()) != hipSuccess) {
                wfail(wi, "entanglement: worker stream failed (%d)", 0);
                return;
            }
            double gms = 0.0, ems = 0.0;
            for (size_t j = 0; j + 2 < tev.size(); j += 3) {
                float a = 0.0f, c = 0.0f;
                (void)hipEventElapsedTime(&a, tev[j], tev[j + 1]);
                (void)hipEventElapsedTime(&c, tev[j + 1], tev[j + 2]);
                gms += a;
                ems += c;
            }
            std::lock_guard<std::mutex> lk(g_stats_mu);
            g_stats.gram_ms += gms;
            g_stats.gram_flops += flops;
            g_stats.eig_ms += ems;
        };
        if (nw > 1 && !pool->warm[cplx ? 1 : 0][want_sigma ? 1 : 0]) {
            // Before the first worker threads on this device: rocBLAS's kernels are loaded and each
            // rocSOLVER routine the workers call is run once on this thread, so that no kernel is
            // loaded by several threads at once on their first calls.
            blas_initialize_once();
            const int64_t kw = std::min<int64_t>(maxk, 2048);
            hipLaunchKernelGGL(k_warm_matrix, dim3((unsigned)((kw * kw + 255) / 256)), dim3(256), 0, st, kw,
                               cplx ? 1 : 0, dR);
            // (its values go to dSall and its info to the second half of dinfo: dWall and the first
            // half may already hold the results of the large blocks solved above)
            rocblas_status rs = cplx ? rocsolver_zheevd(h, rocblas_evect_none, rocblas_fill_upper, (rocblas_int)kw,
                                                        (rocblas_double_complex *)dR, (rocblas_int)kw, dSall, dE,
                                                        dinfo + njobs)
                                     : rocsolver_dsyevd(h, rocblas_evect_none, rocblas_fill_upper, (rocblas_int)kw, dR,
                                                        (rocblas_int)kw, dSall, dE, dinfo + njobs);
            if (rs == rocblas_status_success && want_sigma)
                rs = cplx ? rocsolver_zgesvd(h, rocblas_svect_none, rocblas_svect_none, (rocblas_int)kw,
                                             (rocblas_int)kw, (rocblas_double_complex *)dR, (rocblas_int)kw, dSall,
                                             nullptr, 1, nullptr, 1, dE, rocblas_outofplace, dinfo + njobs)
                          : rocsolver_dgesvd(h, rocblas_svect_none, rocblas_svect_none, (rocblas_int)kw,
                                             (rocblas_int)kw, dR, (rocblas_int)kw, dSall, nullptr, 1, nullptr, 1, dE,
                                             rocblas_outofplace, dinfo + njobs);
            if (rs != rocblas_status_success || hipStreamSynchronize(st) != hipSuccess) {
                s = fail(ED_EHIP, "entanglement: solver warm-up failed (%d)", (int)rs);
                break;
            }
            pool->warm[cplx ? 1 : 0][want_sigma ? 1 : 0] = true;
        }
        if (nw == 1) {
            worker(0, st, h);
        } else {
            std::vector<std::thread> th;
            for (int wi = 0; wi < nw; ++wi) th.emplace_back(worker, wi, pool->st[wi], pool->h[wi]);
            for (std::thread &t : th) t.join();
            for (int wi = 0; wi < nw; ++wi)
                if (hipStreamSynchronize(pool->st[wi]) != hipSuccess && wmsg[wi].empty())
                    wfail(wi, "entanglement: worker stream failed (%d)", 0);
        }
        for (int wi = 0; wi < nw && s == ED_OK; ++wi)
            if (!wmsg[wi].empty()) s = fail(ED_EHIP, "%s", wmsg[wi].c_str());
        if (s) break;
        std::vector<double> wv((size_t)nvals), sv(want_sigma ? (size_t)nvals : 0);
        std::vector<rocblas_int> info((size_t)(2 * njobs), 0);
        if (hipMemcpyAsync(wv.data(), dWall, sizeof(double) * (size_t)nvals, hipMemcpyDeviceToHost, st) ||
            (want_sigma &&
             hipMemcpyAsync(sv.data(), dSall, sizeof(double) * (size_t)nvals, hipMemcpyDeviceToHost, st)) ||
            hipMemcpyAsync(info.data(), dinfo, sizeof(rocblas_int) * (size_t)(want_sigma ? 2 * njobs : njobs),
                           hipMemcpyDeviceToHost, st) ||
            hipStreamSynchronize(st)) {
            s = fail(ED_EHIP, "entanglement: readback failed");
            break;
        }
        for (int b = 0; b < njobs && s == ED_OK; ++b) {
            if (info[b] != 0) s = fail(ED_EHIP, "entanglement: syevd info %d", (int)info[b]);
            else if (want_sigma && info[njobs + b] != 0)
                s = fail(ED_EHIP, "entanglement: gesvd info %d", (int)info[njobs + b]);
        }
        if (s) break;
        // block order, as the blocks are listed in the plan
        for (double v : wv) lambda.push_back(v > 0.0 ? v : 0.0);
        for (double v : sv) sigma.push_back(v);
    } while (0);
    return s;
}

} // namespace ed

using namespace ed;

extern "C" int ed_entanglement_entropy(int N, int Q_up, int Q_dn, int64_t maskA, int complex_vectors,
                                       const void *d_psi, double *out, int64_t *nsv_out, void *stream)
{
    if (!out) return fail(ED_EINVAL, "ed_entanglement_entropy: null output");
    std::vector<double> lam, sig;
    ED_TRY(block_spectra(N, Q_up, Q_dn, maskA, complex_vectors != 0, d_psi, stream_of(stream), lam, sig, true));
    double svn = 0.0, sref = 0.0;
    int64_t nsv = 0;
    for (double l : lam)
        if (l > 0.0) svn -= l * std::log(l);
    for (double s : sig)
        if (s != 0.0) {
            sref -= s * std::log(s);
            ++nsv;
        }
    out[0] = svn;
    out[1] = sref;
    if (nsv_out) *nsv_out = nsv;
    return ED_OK;
}

extern "C" int ed_entanglement_spectrum(int N, int Q_up, int Q_dn, int64_t maskA, int complex_vectors,
                                        const void *d_psi, double *lambda_out, int64_t cap, int64_t *nlambda,
                                        void *stream)
{
    std::vector<double> lam, sig;
    ED_TRY(block_spectra(N, Q_up, Q_dn, maskA, complex_vectors != 0, d_psi, stream_of(stream), lam, sig, false));
    std::sort(lam.begin(), lam.end(), [](double a, double b) { return a > b; });
    if (nlambda) *nlambda = (int64_t)lam.size();
    if (lambda_out)
        for (int64_t i = 0; i < std::min<int64_t>(cap, (int64_t)lam.size()); ++i) lambda_out[i] = lam[(size_t)i];
    return ED_OK;
}

extern "C" int ed_entanglement_entropy_vn(int N, int Q_up, int Q_dn, int64_t maskA, int complex_vectors,
                                          const void *d_psi, double *s_vn, void *stream)
{
    if (!s_vn) return fail(ED_EINVAL, "ed_entanglement_entropy_vn: null output");
    std::vector<double> lam, sig;
    ED_TRY(block_spectra(N, Q_up, Q_dn, maskA, complex_vectors != 0, d_psi, stream_of(stream), lam, sig, false));
    double svn = 0.0;
    for (double l : lam)
        if (l > 0.0) svn -= l * std::log(l);
    *s_vn = svn;
    return ED_OK;
}

// S_vN of many states of one sector and cut: the blocks of rho_A have the same sizes in every
// state, so each block size is solved once for all states of a chunk (one Gram launch over the
// chunk, one strided-batched rocSOLVER dsyevd/zheevd) instead of once per state and block.  The
// per-state arithmetic is that of ed_entanglement_entropy_vn: the same gather, Gram kernel and
// block order of the sum; only the eigenvalue routine is rocSOLVER's batched variant.
extern "C" int ed_entanglement_entropy_vn_batch(int N, int Q_up, int Q_dn, int64_t maskA, int complex_vectors,
                                                int nstates, const void *const *d_psis, double *s_vn, void *stream)
{
    if (!s_vn || !d_psis || nstates < 1) return fail(ED_EINVAL, "ed_entanglement_entropy_vn_batch: bad arguments");
    if (N < 2 || N > 31 || Q_up < 0 || Q_dn < 0 || Q_up > N || Q_dn > N)
        return fail(ED_EINVAL, "ed_entanglement_entropy_vn_batch: bad sector");
    for (int i = 0; i < nstates; ++i)
        if (!d_psis[i]) return fail(ED_EINVAL, "ed_entanglement_entropy_vn_batch: null state %d", i);
    const uint32_t full = (1u << N) - 1u, mA = (uint32_t)maskA & full;
    if (mA == 0 || mA == full) return fail(ED_EINVAL, "entanglement: A must be a proper subset of the sites");
    const bool cplx = complex_vectors != 0;
    const int NC = cplx ? 2 : 1;
    hipStream_t st = stream_of(stream);
    Plan P;
    ED_TRY(make_plan(N, Q_up, Q_dn, mA, P));
    const int64_t dim_up = (int64_t)P.cu.size() / 3, dim_dn = (int64_t)P.cd.size() / 3;
    std::vector<int> jobs;
    std::vector<int64_t> woff;
    int64_t maxk = 1, nvals = 0, rtot = 0; // rtot: doubles of every block's rho, packed
    std::vector<int64_t> roff;
    for (int g = 0; g < (int)P.geo.size(); ++g) {
        const BlockGeo &G = P.geo[g];
        if (G.rows == 0 || G.cols == 0) continue;
        maxk = std::max(maxk, G.rows);
        jobs.push_back(g);
        woff.push_back(nvals);
        roff.push_back(rtot);
        nvals += G.rows;
        rtot += NC * G.rows * G.rows;
    }
    const int njobs = (int)jobs.size();
    rocblas_handle h = solver_handle();
    if (!h) return fail(ED_EHIP, "entanglement: rocblas handle creation failed");
    rocblas_set_stream(h, st);
    // states per chunk: every block matrix and rho of a chunk resident, within half the free memory
    size_t free_b = 0, total_b = 0;
    ED_HIP(hipMemGetInfo(&free_b, &total_b));
    const double per_state = 8.0 * ((double)NC * (double)P.total + (double)rtot + 2.0 * (double)nvals + (double)maxk) +
                             sizeof(rocblas_int) * (double)njobs;
    const int sb = (int)std::max<double>(1.0, std::min<double>((double)nstates, 0.5 * (double)free_b / per_state));
    static thread_local Arena t_eeb_arena;
    Scratch S(t_eeb_arena, st);
    const size_t bM = Scratch::round(sizeof(double) * NC * (size_t)P.total * sb),
                 bR = Scratch::round(sizeof(double) * (size_t)rtot * sb),
                 bW = Scratch::round(sizeof(double) * (size_t)nvals * sb),
                 bE = Scratch::round(sizeof(double) * (size_t)maxk * sb),
                 bU = Scratch::round(sizeof(int32_t) * P.cu.size()), bD = Scratch::round(sizeof(int32_t) * P.cd.size()),
                 bG = Scratch::round(sizeof(BlockGeo) * P.geo.size()),
                 bI = Scratch::round(sizeof(rocblas_int) * (size_t)njobs * sb);
    ED_TRY(S.reserve(bM + bR + bW + bE + bU + bD + bG + bI));
    double *dM = (double *)S.take(bM), *dR = (double *)S.take(bR), *dW = (double *)S.take(bW),
           *dE = (double *)S.take(bE);
    int32_t *dcu = (int32_t *)S.take(bU), *dcd = (int32_t *)S.take(bD);
    BlockGeo *dgeo = (BlockGeo *)S.take(bG);
    rocblas_int *dinfo = (rocblas_int *)S.take(bI);
    if (hipMemcpyAsync(dcu, P.cu.data(), sizeof(int32_t) * P.cu.size(), hipMemcpyHostToDevice, st) ||
        hipMemcpyAsync(dcd, P.cd.data(), sizeof(int32_t) * P.cd.size(), hipMemcpyHostToDevice, st) ||
        hipMemcpyAsync(dgeo, P.geo.data(), sizeof(BlockGeo) * P.geo.size(), hipMemcpyHostToDevice, st))
        return fail(ED_EHIP, "entanglement: upload failed");
    const int64_t D = dim_up * dim_dn;
    const unsigned grid = (unsigned)((D + 255) / 256);
    std::vector<double> wv((size_t)nvals * sb);
    std::vector<rocblas_int> info((size_t)njobs * sb);
    for (int c0 = 0; c0 < nstates; c0 += sb) {
        const int nc = std::min(sb, nstates - c0);
        for (int i = 0; i < nc; ++i) {
            double *Mi = dM + (size_t)NC * (size_t)P.total * i;
            if (cplx)
                hipLaunchKernelGGL(k_gather_blocks<true>, dim3(grid), dim3(256), 0, st, dim_up, dim_dn, dcu, dcd,
                                   P.nbd, dgeo, (const double *)d_psis[c0 + i], Mi);
            else
                hipLaunchKernelGGL(k_gather_blocks<false>, dim3(grid), dim3(256), 0, st, dim_up, dim_dn, dcu, dcd,
                                   P.nbd, dgeo, (const double *)d_psis[c0 + i], Mi);
        }
        ED_CHECK_LAUNCH();
        for (int b = 0; b < njobs; ++b) {
            const BlockGeo &G = P.geo[jobs[b]];
            const int64_t k = G.rows, K = G.cols;
            const int ntile = (int)((k + TT - 1) / TT);
            const int64_t nblk = (int64_t)ntile * (ntile + 1) / 2;
            double *rho = dR + (size_t)roff[b] * sb; // nc matrices of NC k^2 doubles
            const int64_t sA = NC * P.total, sR = NC * k * k;
            if (cplx)
                hipLaunchKernelGGL(k_gram_mfma<true>, dim3((unsigned)nblk, (unsigned)nc), dim3(256), 0, st, k, K,
                                   (const double *)(dM + NC * G.off), rho, ntile, sA, sR);
            else
                hipLaunchKernelGGL(k_gram_mfma<false>, dim3((unsigned)nblk, (unsigned)nc), dim3(256), 0, st, k, K,
                                   (const double *)(dM + NC * G.off), rho, ntile, sA, sR);
            ED_CHECK_LAUNCH();
            // eigenvalues of state i's block into dW[i * nvals + woff[b] ...]
            const rocblas_status rs =
                cplx ? rocsolver_zheevd_strided_batched(h, rocblas_evect_none, rocblas_fill_upper, (rocblas_int)k,
                                                        (rocblas_double_complex *)rho, (rocblas_int)k, k * k,
                                                        dW + woff[b], nvals, dE, maxk, dinfo + (size_t)b * sb,
                                                        (rocblas_int)nc)
                     : rocsolver_dsyevd_strided_batched(h, rocblas_evect_none, rocblas_fill_upper, (rocblas_int)k, rho,
                                                        (rocblas_int)k, k * k, dW + woff[b], nvals, dE, maxk,
                                                        dinfo + (size_t)b * sb, (rocblas_int)nc);
            if (rs != rocblas_status_success)
                return fail(ED_EHIP, "entanglement: rocsolver batched syevd failed (%d)", (int)rs);
        }
        ED_HIP(hipMemcpyAsync(wv.data(), dW, sizeof(double) * (size_t)nvals * nc, hipMemcpyDeviceToHost, st));
        ED_HIP(hipMemcpyAsync(info.data(), dinfo, sizeof(rocblas_int) * (size_t)njobs * sb, hipMemcpyDeviceToHost, st));
        ED_HIP(hipStreamSynchronize(st));
        for (int b = 0; b < njobs; ++b)
            for (int i = 0; i < nc; ++i)
                if (info[(size_t)b * sb + i] != 0)
                    return fail(ED_EHIP, "entanglement: batched syevd info %d (state %d)", (int)info[(size_t)b * sb + i],
                                c0 + i);
        for (int i = 0; i < nc; ++i) {
            // block order, as the blocks are listed in the plan (ed_entanglement_entropy_vn's sum)
            double svn = 0.0;
            for (int64_t j = 0; j < nvals; ++j) {
                const double l = wv[(size_t)i * nvals + j];
                if (l > 0.0) svn -= l * std::log(l);
            }
            s_vn[c0 + i] = svn;
        }
    }
    return ED_OK;
}

extern "C" int ed_entanglement_stats(double *gram_ms, double *gram_flops, double *eig_ms, int reset)
{
    std::lock_guard<std::mutex> lk(g_stats_mu);
    if (gram_ms) *gram_ms = g_stats.gram_ms;
    if (gram_flops) *gram_flops = g_stats.gram_flops;
    if (eig_ms) *eig_ms = g_stats.eig_ms;
    if (reset) g_stats = EeStats();
    return ED_OK;
}

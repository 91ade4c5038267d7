// ed_batch.hip -- many independent Lanczos ground-state solves in one stream of launches.
//
// The reference solves one Hubbard instance at a time (Lanczos<T>::eigen,
// MBL_test/include/HubMatrix_GPU.cuh:561-613, called once per sample/sector by the drivers,
// e.g. drude_v2_cal_Kagome_randomHop.cu:129-175).  For sweeps over uneven sectors, twist
// fluxes or disorder realisations that leaves the GPU mostly idle on the small instances and
// pays a host round trip per check on every one.  Here all instances advance together:
//   * their vectors are packed into three shared buffers (the rotation A/B/C of the
//     single-instance driver), each instance at its own offset;
//   * one launch of each kernel covers every tile of every active instance (a tile list
//     with per-instance offsets; instances that have converged skip their tiles);
//   * reductions stay per instance and deterministic (per-tile partials, one finalize
//     block per instance);
//   * the step number and the active flags live in device memory, so ten steps (the
//     reference's convergence-check cadence) are captured once in a hipGraph per buffer
//     rotation phase and replayed; the host only reads a/b at the checks.
// Each instance follows exactly the recurrence and convergence test of ed_lanczos_eigen
// (eigenvalue only; no Ritz vector).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "ed_common.hpp"
#include "ed_hubbard.hpp"
#include "ed_hv_otf_body.hpp"

namespace ed {

constexpr int BW = 16; // pass-1 slab width in the batch (small sectors have narrow rows)
constexpr int BR = 4;  // pass-2 rows per tile
constexpr int BCH = 2048; // reals per update chunk

struct BInst {
    OtfArgs a;
    int64_t voff;          // offset of the instance's vector in doubles
    int64_t nreal;         // doubles in the instance's vector
    int64_t t1_off, t1_n;  // pass-1 tiles
    int64_t t2_off, t2_n;  // pass-2 tiles
    int64_t c_off, c_n;    // update chunks
    int64_t nrp1, ncb2;
    int64_t p_off;         // offset in the partial buffer (max(t2_n, c_n) entries)
    int32_t has_up, pad;
};

// instance owning global tile T: last i with off[i] <= T (offsets non-decreasing)
template <int FIELD>
__device__ __forceinline__ int find_inst(const BInst *__restrict__ in, int n, int64_t T)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int64_t o = FIELD == 1 ? in[mid].t1_off : FIELD == 2 ? in[mid].t2_off : in[mid].c_off;
        if (o <= T) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <typename V, typename TV>
__global__ __launch_bounds__(256) void k_b_up(const BInst *__restrict__ in, int n, int64_t ntiles, int64_t per_xcd,
                                              int maxnops, const int *__restrict__ active,
                                              const double *__restrict__ bdev, int ldb,
                                              const int *__restrict__ dstep, const double *x, double *out)
{
    constexpr int ORB = 256 / BW;
    extern __shared__ double smem_b[];
    TV *s_val = reinterpret_cast<TV *>(smem_b);
    int32_t *s_col = reinterpret_cast<int32_t *>(s_val + ORB * maxnops);
    __shared__ int s_cnt[ORB];
    const int64_t b = blockIdx.x;
    const int64_t T = (b & 7) * per_xcd + (b >> 3);
    if (T >= ntiles) return;
    const int i = find_inst<1>(in, n, T);
    if (!active[i] || !in[i].has_up) return;
    const BInst &I = in[i];
    const int step = *dstep;
    const double sc = 1.0 / bdev[(int64_t)i * ldb + step];
    otf_up_body<V, TV, true, BW>(I.a, T - I.t1_off, I.nrp1, reinterpret_cast<const V *>(x + I.voff),
                                 reinterpret_cast<V *>(out + I.voff), sc, s_val, s_col, s_cnt);
}

template <typename V, typename TV>
__global__ __launch_bounds__(256) void k_b_dn(const BInst *__restrict__ in, int n, int64_t ntiles, int64_t per_xcd,
                                              const int *__restrict__ active, const double *__restrict__ bdev,
                                              int ldb, const int *__restrict__ dstep, const double *x, double *ypart,
                                              double *y, double *partial)
{
    __shared__ double s_red[4];
    const int64_t b = blockIdx.x;
    const int64_t T = (b & 7) * per_xcd + (b >> 3);
    if (T >= ntiles) return;
    const int i = find_inst<2>(in, n, T);
    if (!active[i]) return;
    const BInst &I = in[i];
    const int step = *dstep;
    const double bN = bdev[(int64_t)i * ldb + step];
    const double sc = 1.0 / bN;
    const V *xv = reinterpret_cast<const V *>(x + I.voff);
    V *yp = reinterpret_cast<V *>(ypart + I.voff);
    V *yy = reinterpret_cast<V *>(y + I.voff);
    double r;
    // ypart (the UP partial) is also the v output; y holds v_{N-1} on entry (w)
    if (step == 0)
        r = otf_dn_body<V, TV, BR, HV_SCALE | HV_WRITEV | HV_DOT, true>(I.a, T - I.t2_off, I.ncb2, I.has_up != 0, xv, yp,
                                                                 yy, yp, yy, sc, 0.0, s_red);
    else
        r = otf_dn_body<V, TV, BR, HV_SCALE | HV_WRITEV | HV_AXPY | HV_DOT, true>(I.a, T - I.t2_off, I.ncb2,
                                                                           I.has_up != 0, xv, yp, yy, yp, yy, sc,
                                                                           -bN, s_red);
    if (threadIdx.x == 0) partial[I.p_off + (T - I.t2_off)] = r;
}

// one block per instance: dst[i*ld + step(+1)] = sum (mode 0) or sqrt(sum) (mode 1) of its
// partials, in the fixed order of the single-instance finalize
__global__ __launch_bounds__(1024) void k_b_finalize(const BInst *__restrict__ in, const int *__restrict__ active,
                                                     const double *__restrict__ partial, int which, double *dst,
                                                     int ld, const int *dstep)
{
    __shared__ double red[16];
    const int i = blockIdx.x;
    const int step = *dstep;
    if (active[i]) {
        const BInst &I = in[i];
        const int64_t np = which == 0 ? I.t2_n : I.c_n;
        double s = 0.0;
        for (int64_t k = threadIdx.x; k < np; k += 1024) s += partial[I.p_off + k];
        s = block_sum<1024>(s, red);
        if (threadIdx.x == 0) {
            if (which == 0) dst[(int64_t)i * ld + step] = s;
            else dst[(int64_t)i * ld + step + 1] = sqrt(s);
        }
    }
}

// last kernel of a step: the step counter lives in device memory so captured steps replay
__global__ void k_b_bump(int *dstep) { dstep[0] += 1; }

// f = -a v + f ; partial |f|^2 per 2048-real chunk
__global__ __launch_bounds__(256) void k_b_update(const BInst *__restrict__ in, int n, int64_t ntiles,
                                                  const int *__restrict__ active, const double *__restrict__ adev,
                                                  int lda, const int *__restrict__ dstep, const double *v, double *f,
                                                  double *partial)
{
    __shared__ double red[4];
    const int64_t T = blockIdx.x;
    if (T >= ntiles) return;
    const int i = find_inst<3>(in, n, T);
    if (!active[i]) return;
    const BInst &I = in[i];
    const double c = -adev[(int64_t)i * lda + *dstep];
    const int64_t ch = T - I.c_off;
    const int64_t base = I.voff + ch * BCH + threadIdx.x, end = I.voff + I.nreal;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < BCH / 256; ++k) {
        const int64_t e = base + (int64_t)k * 256;
        if (e < end) {
            const double r = fma(c, v[e], f[e]);
            f[e] = r;
            s = fma(r, r, s);
        }
    }
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) partial[I.p_off + ch] = s;
}

struct Batch {
    int n = 0;
    bool cvec = false, ctab = false;
    std::vector<const Hubbard *> ops;
    std::vector<BInst> host;
    BInst *dinst = nullptr;
    int64_t total_real = 0, t1 = 0, t2 = 0, tc = 0, pcap = 0;
    int maxnops = 1;
    ~Batch() { dfree(dinst); }
};

} // namespace ed

using namespace ed;

extern "C" int ed_batch_create(int ninst, const ed_hubbard *ops, int complex_vectors, ed_batch *out)
{
    if (!out || ninst < 1 || !ops) return fail(ED_EINVAL, "ed_batch_create: bad arguments");
    *out = nullptr;
    Batch *B = new Batch();
    B->n = ninst;
    B->cvec = complex_vectors != 0;
    int64_t voff = 0, t1 = 0, t2 = 0, tc = 0, poff = 0;
    for (int i = 0; i < ninst; ++i) {
        const Hubbard *h = reinterpret_cast<const Hubbard *>(ops[i]);
        if (!h || h->algo != 3) {
            delete B;
            return fail(ED_EINVAL, "ed_batch_create: instance %d must come from ed_hubbard_create_ops", i);
        }
        if (i == 0) B->ctab = h->cplx_tables;
        if (h->cplx_tables != B->ctab || (h->cplx_tables && !B->cvec)) {
            delete B;
            return fail(ED_EINVAL, "ed_batch_create: instances must share the table type (complex needs complex vectors)");
        }
        B->ops.push_back(h);
        BInst I{};
        I.a = h->otf;
        const int64_t du = h->otf.dim_up, dd = h->otf.dim_dn;
        I.nreal = du * dd * (B->cvec ? 2 : 1);
        I.voff = voff;
        voff += (I.nreal + 31) / 32 * 32; // 256 B aligned instance vectors
        I.has_up = h->otf.up.nops > 0 ? 1 : 0;
        I.nrp1 = (du + 256 / BW - 1) / (256 / BW);
        I.t1_n = ((dd + BW - 1) / BW) * I.nrp1;
        I.t1_off = t1;
        t1 += I.t1_n;
        I.ncb2 = (dd + 255) / 256;
        I.t2_n = ((du + BR - 1) / BR) * I.ncb2;
        I.t2_off = t2;
        t2 += I.t2_n;
        I.c_n = (I.nreal + BCH - 1) / BCH;
        I.c_off = tc;
        tc += I.c_n;
        I.p_off = poff;
        poff += std::max(I.t2_n, I.c_n);
        B->maxnops = std::max(B->maxnops, h->otf.up.nops);
        B->host.push_back(I);
    }
    B->total_real = voff;
    B->t1 = t1;
    B->t2 = t2;
    B->tc = tc;
    B->pcap = poff;
    if (int s = dmalloc((void **)&B->dinst, sizeof(BInst) * ninst)) {
        delete B;
        return s;
    }
    if (hipMemcpy(B->dinst, B->host.data(), sizeof(BInst) * ninst, hipMemcpyHostToDevice) != hipSuccess) {
        delete B;
        return fail(ED_EHIP, "ed_batch_create: upload failed");
    }
    *out = reinterpret_cast<ed_batch>(B);
    return ED_OK;
}

extern "C" int ed_batch_destroy(ed_batch bb)
{
    delete reinterpret_cast<Batch *>(bb);
    return ED_OK;
}

template <typename V, typename TV>
static int enqueue_step(const Batch &B, double *A, double *Bb, double *C, double *partial, const int *active,
                        double *adev, double *bdev, int ld, int *dstep, hipStream_t st)
{
    const int64_t per1 = (B.t1 + 7) / 8, per2 = (B.t2 + 7) / 8;
    const size_t sh = (size_t)(256 / BW) * B.maxnops * (sizeof(TV) + sizeof(int32_t));
    // pass 1: UP partial of every active instance -> C (the v buffer)
    hipLaunchKernelGGL((k_b_up<V, TV>), dim3((unsigned)(per1 * 8)), dim3(256), sh, st, B.dinst, B.n, B.t1, per1,
                       B.maxnops, active, bdev, ld, dstep, A, C);
    // pass 2: B = H v - b B, C = v, partial <v, B>
    hipLaunchKernelGGL((k_b_dn<V, TV>), dim3((unsigned)(per2 * 8)), dim3(256), 0, st, B.dinst, B.n, B.t2, per2,
                       active, bdev, ld, dstep, A, C, Bb, partial);
    hipLaunchKernelGGL(k_b_finalize, dim3(B.n), dim3(1024), 0, st, B.dinst, active, partial, 0, adev, ld, dstep);
    hipLaunchKernelGGL(k_b_update, dim3((unsigned)B.tc), dim3(256), 0, st, B.dinst, B.n, B.tc, active, adev, ld,
                       dstep, C, Bb, partial);
    hipLaunchKernelGGL(k_b_finalize, dim3(B.n), dim3(1024), 0, st, B.dinst, active, partial, 1, bdev, ld, dstep);
    hipLaunchKernelGGL(k_b_bump, dim3(1), dim3(1), 0, st, dstep);
    ED_CHECK_LAUNCH();
    return ED_OK;
}

extern "C" int ed_batch_lowest(ed_batch bb, const ed_lanczos_cfg *cfg, uint32_t seed, int use_graph, double *evals,
                               int *nsteps, int *status, void *stream)
{
    Batch *B = reinterpret_cast<Batch *>(bb);
    if (!B || !evals) return fail(ED_EINVAL, "ed_batch_lowest: bad arguments");
    const int maxN = cfg && cfg->max_steps > 0 ? cfg->max_steps : 1000;
    const double tol = cfg && cfg->tol > 0 ? cfg->tol : 0.01 * std::sqrt(std::numeric_limits<double>::epsilon());
    const int every = 10;
    if (cfg && cfg->check_every > 0 && cfg->check_every != 10)
        return fail(ED_EUNSUPPORTED, "ed_batch_lowest: the batch checks every 10 steps");
    if (cfg && cfg->reorth) return fail(ED_EUNSUPPORTED, "ed_batch_lowest: no reorthogonalisation in the batch");
    // work on a private stream (graph capture needs a non-default stream), ordered after
    // the caller's stream; synchronised before returning
    hipStream_t cst = stream_of(stream), st = nullptr;
    hipEvent_t ev = nullptr;
    ED_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess || hipEventRecord(ev, cst) != hipSuccess ||
        hipStreamWaitEvent(st, ev, 0) != hipSuccess) {
        (void)hipStreamDestroy(st);
        if (ev) (void)hipEventDestroy(ev);
        return fail(ED_EHIP, "ed_batch_lowest: stream setup failed");
    }
    stream = st;
    const int n = B->n, ld = maxN + 3;
    double *buf[3] = {nullptr, nullptr, nullptr}, *partial = nullptr, *adev = nullptr, *bdev = nullptr;
    int *active = nullptr, *dstep = nullptr;
    hipGraphExec_t gexec[3] = {nullptr, nullptr, nullptr};
    hipGraph_t graph[3] = {nullptr, nullptr, nullptr};
    int s = ED_OK;
    auto cleanup = [&] {
        for (int k = 0; k < 3; ++k) {
            if (gexec[k]) (void)hipGraphExecDestroy(gexec[k]);
            if (graph[k]) (void)hipGraphDestroy(graph[k]);
            dfree(buf[k]);
        }
        dfree(partial);
        dfree(adev);
        dfree(bdev);
        dfree(active);
        dfree(dstep);
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
        (void)hipEventDestroy(ev);
    };
    do {
        for (int k = 0; k < 3 && !s; ++k) s = dmalloc((void **)&buf[k], sizeof(double) * (size_t)B->total_real);
        if (s) break;
        if ((s = dmalloc((void **)&partial, sizeof(double) * (size_t)std::max<int64_t>(B->pcap, 1))) ||
            (s = dmalloc((void **)&adev, sizeof(double) * (size_t)n * ld)) ||
            (s = dmalloc((void **)&bdev, sizeof(double) * (size_t)n * ld)) ||
            (s = dmalloc((void **)&active, sizeof(int) * n)) || (s = dmalloc((void **)&dstep, sizeof(int) * 2)))
            break;
        // start vectors (the single-instance GenRand vector of each instance's length) and b_0
        std::vector<double> hb0(n);
        for (int i = 0; i < n && !s; ++i) {
            const BInst &I = B->host[i];
            const int64_t len = I.nreal / (B->cvec ? 2 : 1);
            s = ed_random_vector(len, seed, B->cvec, buf[0] + I.voff, 1, stream);
            if (!s) {
                double out2[2];
                s = ed_dot(len, B->cvec, buf[0] + I.voff, buf[0] + I.voff, out2, stream);
                hb0[i] = std::sqrt(out2[0]);
            }
        }
        if (s) break;
        std::vector<double> hb(n * (size_t)ld, 0.0), ha(n * (size_t)ld, 0.0);
        for (int i = 0; i < n; ++i) hb[(size_t)i * ld] = hb0[i];
        std::vector<int> hact(n, 1), hstep2 = {0, 0};
        if (hipStreamSynchronize(st) != hipSuccess) {
            s = fail(ED_EHIP, "ed_batch_lowest: start vectors failed");
            break;
        }
        if (hipMemcpyAsync(bdev, hb.data(), sizeof(double) * n * ld, hipMemcpyHostToDevice, st) ||
            hipMemcpyAsync(active, hact.data(), sizeof(int) * n, hipMemcpyHostToDevice, st) ||
            hipMemcpyAsync(dstep, hstep2.data(), sizeof(int) * 2, hipMemcpyHostToDevice, st)) {
            s = fail(ED_EHIP, "ed_batch_lowest: upload failed");
            break;
        }
        std::vector<int> last(n, -1), done(n, 0), conv(n, 0), processed(n, 0);
        std::vector<double> anorm(n, 0.0), z;
        int remaining = n, phase = 0;
        auto enqueue = [&](int ph, int nstep) -> int {
            for (int k = 0; k < nstep; ++k) {
                double *A = buf[(ph + k) % 3], *Bq = buf[(ph + k + 1) % 3], *Cq = buf[(ph + k + 2) % 3];
                int r = B->cvec ? (B->ctab ? enqueue_step<cplx, cplx>(*B, A, Bq, Cq, partial, active, adev, bdev, ld,
                                                                     dstep, st)
                                           : enqueue_step<cplx, double>(*B, A, Bq, Cq, partial, active, adev, bdev,
                                                                       ld, dstep, st))
                                : enqueue_step<double, double>(*B, A, Bq, Cq, partial, active, adev, bdev, ld, dstep,
                                                               st);
                if (r) return r;
            }
            return ED_OK;
        };
        int N = 0;
        while (remaining > 0 && N < maxN) {
            const int chunk = std::min(every, maxN - N);
            if (use_graph && chunk == every) {
                if (!gexec[phase]) {
                    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                        s = fail(ED_EHIP, "ed_batch_lowest: capture begin failed");
                        break;
                    }
                    int r = enqueue(phase, chunk);
                    hipError_t e = hipStreamEndCapture(st, &graph[phase]);
                    if (r || e != hipSuccess ||
                        hipGraphInstantiate(&gexec[phase], graph[phase], nullptr, nullptr, 0) != hipSuccess) {
                        s = r ? r : fail(ED_EHIP, "ed_batch_lowest: graph capture failed");
                        break;
                    }
                }
                if (hipGraphLaunch(gexec[phase], st) != hipSuccess) {
                    s = fail(ED_EHIP, "ed_batch_lowest: graph launch failed");
                    break;
                }
            } else if ((s = enqueue(phase, chunk))) {
                break;
            }
            N += chunk;
            phase = (phase + chunk) % 3;
            // convergence check of every active instance (same test as ed_lanczos_eigen)
            if (hipMemcpyAsync(ha.data(), adev, sizeof(double) * n * ld, hipMemcpyDeviceToHost, st) ||
                hipMemcpyAsync(hb.data(), bdev, sizeof(double) * n * ld, hipMemcpyDeviceToHost, st) ||
                hipStreamSynchronize(st)) {
                s = fail(ED_EHIP, "ed_batch_lowest: readback failed");
                break;
            }
            bool changed = false;
            for (int i = 0; i < n; ++i) {
                if (done[i]) continue;
                const double *a = &ha[(size_t)i * ld], *b = &hb[(size_t)i * ld];
                const int Nl = N - 1;
                int M = Nl;
                for (int k = processed[i]; k <= Nl; ++k)
                    if (!(b[k + 1] > 0.0) || !std::isfinite(a[k])) {
                        M = k;
                        break;
                    }
                for (int k = processed[i]; k <= M; ++k) {
                    if (k == 0) anorm[i] = std::fabs(a[0] + b[1]);
                    else anorm[i] = std::fmax(anorm[i], b[k] + std::fabs(a[k] + b[k + 1]));
                }
                processed[i] = M + 1;
                z.assign(M + 1, 0.0);
                double lam;
                ed_tridiag_lowest(M + 1, a, b + 1, &lam, z.data());
                evals[i] = lam;
                last[i] = M;
                bool stop = false;
                if (M < Nl || !(b[M + 1] > 0.0)) stop = true;
                else if (std::fabs(b[M + 1] * z[M]) < tol * anorm[i] && std::fabs(z[0]) > tol * anorm[i]) stop = true;
                if (stop) {
                    done[i] = conv[i] = 1;
                    hact[i] = 0;
                    --remaining;
                    changed = true;
                }
            }
            if (changed && hipMemcpyAsync(active, hact.data(), sizeof(int) * n, hipMemcpyHostToDevice, st)) {
                s = fail(ED_EHIP, "ed_batch_lowest: active upload failed");
                break;
            }
        }
        if (s) break;
        for (int i = 0; i < n; ++i) {
            if (nsteps) nsteps[i] = last[i] + 1;
            if (status) status[i] = conv[i] ? ED_OK : ED_ENOCONV;
        }
        if (hipStreamSynchronize(st) != hipSuccess) s = fail(ED_EHIP, "ed_batch_lowest: final sync failed");
    } while (0);
    cleanup();
    return s;
}

// ed_basis.hip -- basis enumeration/ranking and ELL operator tables on the GPU.
//
// Replaces the host loops of H_TBq (MBL_test/include/fermi_operator.h:42-242):
//   * Labeling (44-53): the sector basis is the ascending sequence of N-bit integers
//     with popcount Q.  That order is the colex order, so state[r] is the colex unrank
//     of r (one thread per rank, binomial table in LDS) and address[state[r]] = r.
//   * stacking + Make_H/Make_J/Make_N/Insert_1Fop/Insert_2Fop/Insert_2Nop (64-225): every row of the
//     column-major ELL table is independent and sees the operator calls in call order,
//     so one thread owns one row, replays the op list and drops each element into the
//     first slot whose value is exactly 0.0 -- the same slot the reference picks.
// Both are integer/byte work, bit-exact against the reference's own fermi_operator.h.
#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "ed_common.hpp"
#include "ed_thread.hpp"

namespace ed {

// Set when this library's static objects are destroyed (process exit).  libedhip depends on the
// HIP runtime, so its statics are destroyed before the runtime's: until the flag is set, HIP calls
// from thread-exit destructors are safe.
static std::atomic<bool> g_exiting{false};
namespace {
struct ExitFlag {
    ~ExitFlag() { g_exiting.store(true); }
} g_exit_flag;

struct ThreadBlas {
    rocblas_handle h = nullptr;
    bool tried = false;
    ~ThreadBlas()
    {
        if (h && !process_exiting()) (void)rocblas_destroy_handle(h);
    }
};
thread_local ThreadBlas t_blas;
} // namespace

bool process_exiting() { return g_exiting.load(); }

rocblas_handle thread_blas_handle()
{
    if (!t_blas.tried) {
        t_blas.tried = true;
        if (rocblas_create_handle(&t_blas.h) != rocblas_status_success) t_blas.h = nullptr;
    }
    return t_blas.h;
}

void blas_initialize_once()
{
    static std::mutex mu;
    static bool done[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return;
    std::lock_guard<std::mutex> lk(mu);
    if (done[dev]) return;
    rocblas_initialize();
    done[dev] = true;
}

static thread_local std::string g_err;

void set_error(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

int fail(int status, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return status;
}

const Binom &binom_host()
{
    static Binom b;
    static std::once_flag once;
    std::call_once(once, [] {
        memset(&b, 0, sizeof b);
        for (int n = 0; n < 64; ++n) {
            b.c[n][0] = 1;
            for (int k = 1; k <= n; ++k) b.c[n][k] = b.c[n - 1][k - 1] + (k <= n - 1 ? b.c[n - 1][k] : 0);
        }
    });
    return b;
}

int64_t binom(int n, int k)
{
    if (n < 0 || k < 0 || k > n || n > 63) return 0;
    return binom_host().c[n][k];
}

// Device memory the library allocates itself (dmalloc/dfree, all devices together): current and
// peak bytes, for ed_memory_stats.  rocBLAS/rocSOLVER workspaces are not counted.
namespace {
std::mutex g_mem_mu;
std::unordered_map<void *, size_t> g_mem_size;
int64_t g_mem_cur = 0, g_mem_peak = 0;
} // namespace

int dmalloc(void **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        if (release_idle_thread_pools() > 0) e = hipMalloc(p, bytes);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(ED_ENOMEM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    }
    std::lock_guard<std::mutex> lk(g_mem_mu);
    g_mem_size[*p] = bytes;
    g_mem_cur += (int64_t)bytes;
    if (g_mem_cur > g_mem_peak) g_mem_peak = g_mem_cur;
    return ED_OK;
}

void dfree(void *p)
{
    if (!p) return;
    (void)hipFree(p);
    std::lock_guard<std::mutex> lk(g_mem_mu);
    auto it = g_mem_size.find(p);
    if (it != g_mem_size.end()) {
        g_mem_cur -= (int64_t)it->second;
        g_mem_size.erase(it);
    }
}

Arena::~Arena()
{
    if (base && !process_exiting()) dfree(base);
}

Scratch::~Scratch()
{
    if (cur_) (void)hipStreamSynchronize(st_);
    dfree(own_);
}

int Scratch::reserve(size_t total)
{
    total = std::max<size_t>(total, 256);
    if (total <= kArenaMax) {
        if (a_.cap < total) {
            dfree(a_.base);
            a_.base = nullptr;
            a_.cap = 0;
            ED_TRY(dmalloc((void **)&a_.base, total));
            a_.cap = total;
        }
        cur_ = a_.base;
    } else {
        ED_TRY(dmalloc((void **)&own_, total));
        cur_ = own_;
    }
    end_ = cur_ + total;
    return ED_OK;
}

void *Scratch::take(size_t bytes)
{
    char *p = cur_;
    cur_ += round(bytes);
    return cur_ <= end_ ? p : nullptr;
}

// ---------------------------------------------------------------------------------
// basis: colex unrank.  rank = sum_k C(p_k, k+1) over set-bit positions p_0 < p_1 < ...
// ---------------------------------------------------------------------------------
__global__ void k_basis_unrank(int N, int Q, int64_t count, const int64_t *__restrict__ binom_flat,
                               int32_t *__restrict__ state, int32_t *__restrict__ address)
{
    __shared__ int64_t C[32 * 33];
    for (int t = threadIdx.x; t < 32 * 33; t += blockDim.x) C[t] = binom_flat[(t / 33) * 64 + (t % 33)];
    __syncthreads();
    const int64_t r0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r0 >= count) return;
    int64_t r = r0;
    int32_t s = 0;
    if (Q < 0) {
        s = (int32_t)r0; // Full_Labeling (fermi_operator.h:55-62)
    } else {
        int p = N - 1;
        for (int k = Q; k >= 1; --k) {
            // largest p with C(p, k) <= r
            while (C[p * 33 + k] > r) --p;
            s |= (int32_t)1 << p;
            r -= C[p * 33 + k];
            --p;
        }
    }
    state[r0] = s;
    address[s] = (int32_t)r0;
}

// LOW_MEM::H_TBq::Labeling (TotalSpin/include/fermi_operator_lowmem.h:50-64): prev_permutation
// over occ[0..N-1], starting from Q ones first, visits the occupation patterns in descending
// lexicographic order of (occ[0], occ[1], ...), i.e. descending value of the bit-reversed state.
// Index r is therefore the colex unrank of count-1-r, bit-reversed over N bits.
__global__ void k_basis_unrank_lowmem(int N, int Q, int64_t count, const int64_t *__restrict__ binom_flat,
                                      uint64_t *__restrict__ state)
{
    const int64_t r0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r0 >= count) return;
    int64_t r = count - 1 - r0;
    uint64_t s = 0;
    int p = N - 1;
    for (int k = Q; k >= 1; --k) {
        while (binom_flat[p * 64 + k] > r) --p;
        s |= (uint64_t)1 << (N - 1 - p); // bit p of the reversed value is site N-1-p
        r -= binom_flat[p * 64 + k];
        --p;
    }
    state[r0] = s;
}

} // namespace ed

using namespace ed;

extern "C" int ed_memory_stats(int64_t *current_bytes, int64_t *peak_bytes, int reset_peak)
{
    std::lock_guard<std::mutex> lk(g_mem_mu);
    if (current_bytes) *current_bytes = g_mem_cur;
    if (peak_bytes) *peak_bytes = g_mem_peak;
    if (reset_peak) g_mem_peak = g_mem_cur;
    return ED_OK;
}

extern "C" const char *ed_last_error(void) { return g_err.c_str(); }

extern "C" int ed_version(void) { return 10000; }

extern "C" int ed_device_sync(void)
{
    ED_HIP(hipDeviceSynchronize());
    return ED_OK;
}

extern "C" int ed_basis_count(int N, int Q, int64_t *count)
{
    if (N < 1 || N > 31 || Q > N || !count) return fail(ED_EINVAL, "ed_basis_count: bad N=%d Q=%d", N, Q);
    *count = Q < 0 ? ((int64_t)1 << N) : binom(N, Q);
    return ED_OK;
}

static int64_t *device_binom()
{
    static int64_t *d = nullptr;
    static std::once_flag once;
    std::call_once(once, [] {
        if (hipMalloc(&d, sizeof(Binom)) == hipSuccess) {
            if (hipMemcpy(d, &binom_host(), sizeof(Binom), hipMemcpyHostToDevice) != hipSuccess) {
                (void)hipFree(d);
                d = nullptr;
            }
        } else {
            d = nullptr;
        }
    });
    return d;
}

extern "C" int ed_basis_build(int N, int Q, int32_t *d_state, int32_t *d_address, void *stream)
{
    int64_t count;
    ED_TRY(ed_basis_count(N, Q, &count));
    if (!d_state || !d_address) return fail(ED_EINVAL, "ed_basis_build: null buffer");
    hipStream_t st = stream_of(stream);
    int64_t *db = device_binom();
    if (!db) return fail(ED_ENOMEM, "ed_basis_build: binomial table upload failed");
    // address[] is 0 outside the sector (host_vector::resize value-initialises)
    ED_HIP(hipMemsetAsync(d_address, 0, sizeof(int32_t) * ((size_t)1 << N), st));
    const int B = 256;
    const int64_t grid = (count + B - 1) / B;
    hipLaunchKernelGGL(k_basis_unrank, dim3((unsigned)grid), dim3(B), 0, st, N, Q, count, db, d_state, d_address);
    ED_CHECK_LAUNCH();
    return ED_OK;
}

extern "C" int ed_basis_build_lowmem(int N, int Q, uint64_t *d_state, void *stream)
{
    if (N < 1 || N > 63 || Q < 0 || Q > N || !d_state)
        return fail(ED_EINVAL, "ed_basis_build_lowmem: bad N=%d Q=%d", N, Q);
    const int64_t count = binom(N, Q);
    if (count <= 0 || count > ((int64_t)1 << 31) - 1)
        return fail(ED_EINVAL, "ed_basis_build_lowmem: C(%d,%d) does not fit an int rank", N, Q);
    int64_t *db = device_binom();
    if (!db) return fail(ED_ENOMEM, "ed_basis_build_lowmem: binomial table upload failed");
    const int B = 256;
    hipLaunchKernelGGL(k_basis_unrank_lowmem, dim3((unsigned)((count + B - 1) / B)), dim3(B), 0, stream_of(stream), N,
                       Q, count, db, d_state);
    ED_CHECK_LAUNCH();
    return ED_OK;
}

// ---------------------------------------------------------------------------------
// ELL tables: one thread per row, op list replayed in call order.
// ---------------------------------------------------------------------------------
namespace ed {

struct DevOp {
    int32_t kind, i, j, pad;
    double value, phase;
};

__device__ __forceinline__ int count_between(int64_t n, int i, int j)
{
    // count1_bt_sites (fermi_operator.h:26-40): occupied sites strictly between i and j
    int lo = i < j ? i : j, hi = i < j ? j : i;
    if (hi - lo <= 1) return 0;
    const uint64_t m = (((uint64_t)1 << (hi - lo - 1)) - 1) << (lo + 1);
    return __popcll((uint64_t)n & m);
}

template <bool CPLX>
__global__ void k_tables(int Q, int S, int64_t count, const int32_t *__restrict__ state,
                         const int32_t *__restrict__ address, const DevOp *__restrict__ ops, int nops,
                         double *__restrict__ Ax, int32_t *__restrict__ Aj, int *__restrict__ err)
{
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= count) return;
    const int64_t n = state[a];
    for (int o = 0; o < nops; ++o) {
        const DevOp op = ops[o];
        const int oi = (int)((n >> op.i) & 1), oj = (int)((n >> op.j) & 1);
        bool emit = false;
        int32_t col = 0;
        double vr = 0.0, vi = 0.0, sign = 1.0;
        switch (op.kind) {
        case ED_OP_HOP:
            if (oi != oj) {
                emit = true;
                sign = (count_between(n, op.i, op.j) & 1) ? -1.0 : 1.0;
                col = address[n ^ (((int64_t)1 << op.i) | ((int64_t)1 << op.j))];
                vr = op.value;
            }
            break;
        case ED_OP_HOP_PHASE:
            if (oi != oj) {
                emit = true;
                sign = (count_between(n, op.i, op.j) & 1) ? -1.0 : 1.0;
                col = address[n ^ (((int64_t)1 << op.i) | ((int64_t)1 << op.j))];
                // i occupied, j empty: this row receives c+_i c_j -> t e^{+i phase}
                const double ph = (oi == 1) ? op.phase : -op.phase;
                double s, c;
                sincos(ph, &s, &c);
                vr = op.value * c;
                vi = op.value * s;
            }
            break;
        case ED_OP_CURRENT:
            if (oi != oj) {
                emit = true;
                sign = (count_between(n, op.i, op.j) & 1) ? -1.0 : 1.0;
                col = address[n ^ (((int64_t)1 << op.i) | ((int64_t)1 << op.j))];
                vr = (oi == 0) ? op.value : -op.value; // +t c+_i c_j, -t c+_j c_i
            }
            break;
        case ED_OP_DENSITY:
            if (oi == 1) {
                emit = true;
                col = (int32_t)a;
                vr = op.value;
            }
            break;
        case ED_OP_DENSITY2:
            if (oi == 1 && oj == 1) {
                emit = true;
                col = (int32_t)a;
                vr = op.value;
            }
            break;
        case ED_OP_FERMI1: // Insert_1Fop (fermi_operator.h:151-174): sign from the sites above i
            if (oi == (op.pad & 1)) {
                emit = true;
                sign = (__popcll((uint64_t)n >> (op.i + 1)) & 1) ? -1.0 : 1.0;
                col = address[n ^ ((int64_t)1 << op.i)];
                vr = op.value;
            }
            break;
        case ED_OP_FERMI2: // Insert_2Fop (fermi_operator.h:176-214)
            if (oi == (op.pad & 1) && oj == ((op.pad >> 1) & 1)) {
                emit = true;
                sign = (count_between(n, op.i, op.j) & 1) ? -1.0 : 1.0;
                col = address[n ^ (((int64_t)1 << op.i) | ((int64_t)1 << op.j))];
                vr = op.value;
            }
            break;
        default: break;
        }
        if (!emit) continue;
        int k = 0;
        for (; k < S; ++k) {
            const int64_t idx = (int64_t)k * count + a;
            if (CPLX) {
                if (Ax[2 * idx] != 0.0 || Ax[2 * idx + 1] != 0.0) continue;
                Ax[2 * idx] += sign * vr;
                Ax[2 * idx + 1] += sign * vi;
            } else {
                if (Ax[idx] != 0.0) continue;
                Ax[idx] += sign * vr;
            }
            Aj[idx] = col;
            break;
        }
        if (k == S) {
            atomicOr(err, 1);
            return;
        }
    }
}

} // namespace ed

extern "C" int ed_tables_build(int N, int Q, const int32_t *d_state, const int32_t *d_address, int64_t count,
                               const ed_op *ops, int nops, int complex_tables, double *d_Ax, int32_t *d_Aj,
                               void *stream)
{
    int64_t expect;
    ED_TRY(ed_basis_count(N, Q, &expect));
    if (count != expect) return fail(ED_EINVAL, "ed_tables_build: count %lld != C(N,Q) %lld", (long long)count,
                                     (long long)expect);
    if (nops < 0 || (nops > 0 && !ops) || !d_Ax || !d_Aj || !d_state || !d_address)
        return fail(ED_EINVAL, "ed_tables_build: bad arguments");
    for (int o = 0; o < nops; ++o) {
        const ed_op &op = ops[o];
        const bool one_site = op.kind == ED_OP_DENSITY || op.kind == ED_OP_FERMI1;
        if (op.kind < 0 || op.kind > 6 || op.i < 0 || op.i >= N || (!one_site && (op.j < 0 || op.j >= N)) ||
            (!one_site && op.i == op.j) || ((op.kind == ED_OP_FERMI1 || op.kind == ED_OP_FERMI2) && (op.dag & ~3)))
            return fail(ED_EINVAL, "ed_tables_build: bad op %d (kind %d, i %d, j %d)", o, op.kind, op.i, op.j);
        if (op.kind == ED_OP_HOP_PHASE && !complex_tables)
            return fail(ED_EINVAL, "ed_tables_build: a phase hop needs complex tables");
    }
    const int S = kZSlots * (Q < 0 ? N : Q) + 1;
    hipStream_t st = stream_of(stream);
    const size_t nent = (size_t)count * (size_t)S;
    ED_HIP(hipMemsetAsync(d_Ax, 0, nent * sizeof(double) * (complex_tables ? 2 : 1), st));
    ED_HIP(hipMemsetAsync(d_Aj, 0, nent * sizeof(int32_t), st));
    if (nops == 0) return ED_OK;
    void *dops = nullptr, *derr = nullptr;
    ED_TRY(dmalloc(&dops, sizeof(DevOp) * nops));
    if (int s = dmalloc(&derr, sizeof(int))) {
        dfree(dops);
        return s;
    }
    std::vector<DevOp> hops(nops);
    for (int o = 0; o < nops; ++o) {
        const bool one_site = ops[o].kind == ED_OP_DENSITY || ops[o].kind == ED_OP_FERMI1;
        hops[o] = DevOp{ops[o].kind, ops[o].i, one_site ? ops[o].i : ops[o].j, ops[o].dag, ops[o].value, ops[o].phase};
    }
    int herr = 0;
    int status = ED_OK;
    do {
        if (hipMemcpyAsync(dops, hops.data(), sizeof(DevOp) * nops, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemsetAsync(derr, 0, sizeof(int), st) != hipSuccess) {
            status = fail(ED_EHIP, "ed_tables_build: upload failed");
            break;
        }
        const int B = 256;
        const unsigned grid = (unsigned)((count + B - 1) / B);
        if (complex_tables)
            hipLaunchKernelGGL(k_tables<true>, dim3(grid), dim3(B), 0, st, Q, S, count, d_state, d_address,
                               (const DevOp *)dops, nops, d_Ax, d_Aj, (int *)derr);
        else
            hipLaunchKernelGGL(k_tables<false>, dim3(grid), dim3(B), 0, st, Q, S, count, d_state, d_address,
                               (const DevOp *)dops, nops, d_Ax, d_Aj, (int *)derr);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(&herr, derr, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            status = fail(ED_EHIP, "ed_tables_build: kernel failed");
            break;
        }
        if (herr) status = fail(ED_ESLOT, "ed_tables_build: a row needs more than %d ELL slots", S);
    } while (0);
    dfree(dops);
    dfree(derr);
    return status;
}

// ed_common.hpp -- shared helpers of libedhip (status handling, launch helpers,
// wavefront/block reductions).  Written for gfx950 (64-wide wavefronts).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/edhip.h"

namespace ed {

void set_error(const char *fmt, ...);
int fail(int status, const char *fmt, ...);

#define ED_HIP(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) return ::ed::fail(ED_EHIP, "%s failed: %s (%s:%d)", #call,               \
                                                hipGetErrorString(e_), __FILE__, __LINE__);            \
    } while (0)

#define ED_CHECK_LAUNCH() ED_HIP(hipGetLastError())

#define ED_TRY(call)                                                                                   \
    do {                                                                                               \
        int s_ = (call);                                                                               \
        if (s_ != ED_OK) return s_;                                                                    \
    } while (0)

inline hipStream_t stream_of(void *s) { return reinterpret_cast<hipStream_t>(s); }

constexpr int kWave = 64;
constexpr int kZSlots = 10; // fermi_operator.h:11

// complex128 as a pair; the library never uses std::complex in device code.
struct cplx {
    double re, im;
};

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Deterministic block sum (fixed tree order); result valid in thread 0.
template <int BLOCK>
__device__ __forceinline__ double block_sum(double v, double *lds)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int w = 0; w < BLOCK / kWave; ++w) r += lds[w];
    }
    __syncthreads();
    return r;
}

// binomial coefficients C(n,k), n,k <= 63, computed once on the host
struct Binom {
    int64_t c[64][64];
};
const Binom &binom_host();
int64_t binom(int n, int k);

// device scratch allocation with error reporting; when hipMalloc fails, the calling thread's
// idle grow-only pools (release_idle_thread_pools) are released and the allocation tried once more
int dmalloc(void **p, size_t bytes);
void dfree(void *p);
// Frees the calling thread's grow-only pools that no call of the thread is using (the Lanczos
// basis kept for selective orthogonalisation, which grows to the longest solve: up to max_steps
// vectors); returns the bytes released.
size_t release_idle_thread_pools();

// Grow-only device workspace of one host thread.  hipMalloc and hipFree wait for the whole
// device, so entry points that run many times on concurrent streams (the disorder-ensemble
// solves, eight host threads at once) take their per-call buffers from it instead of allocating
// them; the memory is kept until the thread ends.  Each source file has its own arena, so an
// entry point never shares one with an entry point it calls.
struct Arena {
    char *base = nullptr;
    size_t cap = 0;
    ~Arena();
};
// Per-call view: buffers carved (256 B aligned) from the thread's arena when the call needs at
// most kArenaMax bytes in total, else from one private allocation freed with the view.  The
// view synchronises its stream when it goes out of scope, so no work of this call still uses
// the arena when the thread's next call (maybe on another stream) reuses it.
class Scratch {
  public:
    static constexpr size_t kArenaMax = (size_t)4 << 30;
    Scratch(Arena &a, hipStream_t st) : a_(a), st_(st) {}
    ~Scratch();
    Scratch(const Scratch &) = delete;
    Scratch &operator=(const Scratch &) = delete;
    static size_t round(size_t bytes) { return (bytes + 255) & ~(size_t)255; }
    int reserve(size_t total); // total = sum of round(bytes) of the buffers to take
    void *take(size_t bytes);

  private:
    Arena &a_;
    hipStream_t st_;
    char *own_ = nullptr, *cur_ = nullptr, *end_ = nullptr;
};

} // namespace ed

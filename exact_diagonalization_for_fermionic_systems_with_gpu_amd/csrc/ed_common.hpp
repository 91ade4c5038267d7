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

// device scratch allocation with error reporting
int dmalloc(void **p, size_t bytes);
void dfree(void *p);

} // namespace ed

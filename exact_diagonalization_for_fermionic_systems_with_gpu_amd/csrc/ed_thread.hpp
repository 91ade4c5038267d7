// ed_thread.hpp -- per-host-thread device resources of libedhip (not part of the C ABI).
//
// Policy: every per-thread device resource (rocBLAS handles, grow-only workspaces) is released
// when its thread exits -- a Python pool thread that ends frees what it used.  Once the process
// has begun destroying its static objects (process_exiting()), the release is skipped: the HIP
// runtime may already be gone, and the operating system reclaims the memory.
#pragma once

#include <rocblas/rocblas.h>

namespace ed {

bool process_exiting();

// The calling thread's rocBLAS handle (created on first use, destroyed at thread exit); a handle
// carries its stream, so threads that work on different streams never share one.  nullptr if it
// cannot be created.
rocblas_handle thread_blas_handle();

// rocblas_initialize() once per process and device (it loads rocBLAS's kernels up front), so
// that several host threads never load them concurrently on their first calls.
void blas_initialize_once();

} // namespace ed

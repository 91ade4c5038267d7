// ed_hubbard.hpp -- internal types of the Hubbard operator (not part of the C ABI).
#pragma once

#include "ed_common.hpp"

namespace ed {

struct SpeciesIn {
    int64_t dim;
    int S;
    const double *Ax;
    const int32_t *Aj;
};

// compact form of one species' ELL table
struct Species {
    int64_t dim = 0;
    int64_t nent = 0;
    int max_nnz = 0;
    int64_t *ptr = nullptr; // UP: CSR row pointer [dim+1]; DN: slice pointer [nslices+1]
    int32_t *col = nullptr;
    double *val = nullptr; // double or interleaved complex
    const int32_t *state = nullptr;
};

struct HvArgs {
    int64_t dim_up, dim_dn;
    const int64_t *up_ptr;
    const int32_t *up_col;
    const void *up_val;
    const int64_t *dn_sptr;
    const int32_t *dn_col;
    const void *dn_val;
    const int32_t *state_up;
    const int32_t *state_dn;
    double U;
    int64_t mask;
    const double *Uvec;
    int has_diag;
    int64_t blocks_per_row;
};

struct Hubbard {
    int N = 0;
    bool cplx_tables = false;
    double U = 0.0;
    int64_t mask = 0;
    Species up, dn;
    int32_t *d_state_up = nullptr, *d_state_dn = nullptr;
    double *d_uvec = nullptr;
    HvArgs args{};
    int algo = 2;    // 2: two-pass XCD-aware kernels; 1: single-pass row kernel
    int hv_w = 8;    // pass-1 slab width (doubles)
    int hv_rmax = 8; // largest pass-2 row panel
    void release();
};

int hv_plain(const Hubbard &h, bool cvec, const void *x, void *y, hipStream_t st);
int hv_lanczos(const Hubbard &h, bool cvec, const void *f, void *y, void *v, const void *w, const double *bdev,
               int step, double *partial, hipStream_t st, int64_t *nblocks);
int hv_lanczos_nodot(const Hubbard &h, bool cvec, const void *f, void *y, void *v, const void *w,
                     const double *bdev, int step, hipStream_t st);
int64_t hv_num_blocks(const Hubbard &h);

// deterministic reductions and BLAS-1 (ed_blas.hip)
int finalize_sum(const double *partial, int64_t n, double *dst, int mode, hipStream_t st);
int lanczos_update(int64_t nreal, const double *v, double *f, const double *adev, int step, double *partial,
                   int64_t *nblocks, hipStream_t st);
int ritz_update(int64_t nreal, const double *v, double *f, double *evec, const double *adev, double z, int step,
                hipStream_t st);
int dot_partials(int64_t nreal, const double *a, const double *b, double *partial, int64_t *nblocks,
                 hipStream_t st);
int64_t blas_num_blocks(int64_t nreal);

} // namespace ed

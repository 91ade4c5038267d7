// ed_hubbard.hpp -- internal types of the Hubbard operator (not part of the C ABI).
#pragma once

#include <vector>
#include "ed_common.hpp"

namespace ed {

// fused Lanczos options of the H.v kernels: HV_SCALE (x gathered as x/b[step]), HV_WRITEV
// (v = x/b[step] written), HV_AXPY (y = Hv - b[step] w), HV_DOT (block partials of Re<v,y>)
enum { HV_SCALE = 1, HV_WRITEV = 2, HV_AXPY = 4, HV_DOT = 8 };

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

// operator call of H_TBq, prepared for on-the-fly hop generation (ed_hv_otf.hip)
struct OtfOp {
    int32_t kind, i, j, pad;
    uint32_t m, between;          // 1<<i | 1<<j, sites strictly between i and j
    double vf_re, vf_im;          // value for a row with site i occupied (j empty)
    double vb_re, vb_im;          // value for a row with site j occupied (i empty)
};

struct OtfSpecies {
    int nops = 0, nhop = 0;
    const OtfOp *ops = nullptr;
    OtfOp *ops_own = nullptr;
    const int32_t *addr = nullptr; // 2^N rank table (H_TBq::address)
};

struct OtfArgs {
    int64_t dim_up, dim_dn;
    OtfSpecies up, dn;
    const int32_t *state_up, *state_dn;
    double U;
    int64_t mask;
    const double *Uvec;
    int has_diag;
    const int32_t *up_order; // pass-1 row visiting order (rank list), nullptr: ascending rank
    const uint32_t *up_tab;  // pass-1 entries per row position (column << 8 | value code), or nullptr
    const int32_t *up_cnt;   // entries per row position
};

struct SlabLayout; // ed_hv_slab.hip
void slab_free(SlabLayout *s);

struct Hubbard {
    int N = 0;
    bool cplx_tables = false;
    double U = 0.0;
    int64_t mask = 0;
    Species up, dn;
    int32_t *d_state_up = nullptr, *d_state_dn = nullptr;
    double *d_uvec = nullptr;
    HvArgs args{};
    int algo = 2;    // 3: on-the-fly two-pass; 2: two-pass table kernels; 1: single-pass row kernel
    OtfArgs otf{};
    int32_t *d_addr_up = nullptr, *d_addr_dn = nullptr;
    int hv_w = 8;    // pass-1 slab width (doubles)
    int hv_rmax = 8; // largest pass-2 row panel
    int hv_dnmask = 1; // pass 2 loads only for calls with an element in the column
    int hv_dn2 = 1;    // pass 2 with per-column descriptors cached in LDS over K row panels
    int hv_dnk = 2;
    int hv_upxcd = 1;  // pass 1 tiles partitioned over the 8 XCDs by slab
    int32_t *d_up_order = nullptr; // reverse Cuthill-McKee order of the UP hopping graph
    int hv_uporder = 1;            // pass 1 visits rows in d_up_order (when it exists)
    uint32_t *d_up_tab = nullptr;  // pass-1 entry table in visiting order (dim_up x nops words)
    int32_t *d_up_cnt = nullptr;
    int hv_uptab = 1;              // pass 1 reads d_up_tab (when it exists) instead of generating
    std::vector<OtfOp> h_ops_up, h_ops_dn; // host copies of otf.up/dn.ops (for ed_batch_create)
    std::vector<double> h_uvec;            // host copy of d_uvec (N entries), empty without one
    bool has_fermi = false; // an Insert_1Fop / Insert_2Fop call: table kernels only
    SlabLayout *slab = nullptr; // slab-tiled vector layout (ed_hubbard_slab_layout), or nullptr
    void release();
};

int hv_plain(const Hubbard &h, bool cvec, const void *x, void *y, hipStream_t st);
int hv_lanczos(const Hubbard &h, bool cvec, const void *f, void *y, void *v, const void *w, const double *bdev,
               int step, double *partial, hipStream_t st, int64_t *nblocks);
int hv_lanczos_nodot(const Hubbard &h, bool cvec, const void *f, void *y, void *v, const void *w,
                     const double *bdev, int step, hipStream_t st);
int64_t hv_num_blocks(const Hubbard &h);
int otf_build_up_table(Hubbard &h, hipStream_t st);
template <int FLAGS>
int otf_dispatch(const Hubbard &h, bool cvec, const void *x, void *y, void *vout, const void *w, const double *bdev,
                 int step, double *partial, hipStream_t st, int64_t *nblocks);
int64_t otf_num_blocks(const Hubbard &h);
int otf_prepare(OtfSpecies &sp, const ed_op *ops, int nops, const int32_t *d_address, hipStream_t st);
int up_row_order(int N, int Q, const ed_op *ops, int nops, std::vector<int32_t> &order);
// Lanczos step halves on the slab layout (ed_hv_slab.hip): same contract as hv_lanczos /
// hv_lanczos_nodot on vectors of slab_len() doubles
int64_t slab_len(const Hubbard &h);
// the length a slab layout of this operator has (or would have): dim_up * dim_dn when dim_dn is even
int64_t slab_len_for(const Hubbard &h);
int slab_lanczos(const Hubbard &h, const double *f, double *y, double *v, const double *w, const double *bdev,
                 int step, bool dot, double *partial, hipStream_t st, int64_t *nblocks);
// one step of the Ritz-vector rebuild on the slab layout: slab_lanczos without the dot, with
// ritz_update (f = y - a[step] v, evec += zr v) fused into pass B; y ends as f
int slab_lanczos_ritz(const Hubbard &h, const double *f, double *y, double *v, const double *bdev,
                      const double *adev, double zr, double *evec, int step, hipStream_t st);
int64_t slab_num_partials(const Hubbard &h);
// the same steps keeping no v: the previous f stands for v_{N-1} (scaled by 1 / b[step - 1] where
// it is read) and y holds the UP partial before the result; bit-identical recurrence
int slab_lanczos_vfree(const Hubbard &h, const double *f, const double *f_prev, double *y, const double *bdev,
                       int step, double *partial, hipStream_t st, int64_t *nblocks);
int slab_lanczos_ritz_vfree(const Hubbard &h, const double *f, const double *f_prev, double *y, const double *bdev,
                            const double *adev, double zr, double *evec, int step, hipStream_t st);
// f_next = y - a[step] f / b[step] (in y) and the partials of |f_next|^2 (lanczos_update with v = f / b)
int lanczos_update_vfree(int64_t nreal, const double *f, double *y, const double *adev, const double *bdev, int step,
                         double *partial, int64_t *nblocks, hipStream_t st);
int slab_from_rowmajor(const Hubbard &h, const double *x, double *xl, hipStream_t st);
int slab_to_rowmajor(const Hubbard &h, const double *xl, double *x, hipStream_t st);
// builds the slab layout for a large real sector when it and nvec more layout vectors fit
void slab_auto(Hubbard &h, int nvec);

// deterministic reductions and BLAS-1 (ed_blas.hip)
int finalize_sum(const double *partial, int64_t n, double *dst, int mode, hipStream_t st);
int lanczos_update(int64_t nreal, const double *v, double *f, const double *adev, int step, double *partial,
                   int64_t *nblocks, hipStream_t st);
int ritz_update(int64_t nreal, const double *v, double *f, double *evec, const double *adev, double z, int step,
                hipStream_t st);
int dot_partials(int64_t nreal, const double *a, const double *b, double *partial, int64_t *nblocks,
                 hipStream_t st);
int64_t blas_num_blocks(int64_t nreal);
// The nev lowest eigenvalues of a Lanczos tridiagonal (a[0..M-1], couplings e[0..M-2]) with
// copies of converged eigenvalues merged and spurious values dropped (Cullum-Willoughby), each
// simple one tested |bres s[M-1]| < tol * anorm (ed_lanczos.hip); semi_orth: the recurrence ran
// with selective orthogonalisation.  Returns the number found, *conv = all nev found and passing.
int cw_lowest(int M, const double *a, const double *e, double bres, int nev, double tol, double anorm,
              bool semi_orth, double *out, bool *conv);

} // namespace ed

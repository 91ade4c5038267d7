/*
 * edhip.h -- C ABI of the MI355X exact-diagonalization library (libedhip.so).
 *
 * Plain pointers and sizes only.  "d_" arguments are device pointers (HIP device
 * memory on the current device); everything else is host memory.  Streams are
 * hipStream_t passed as void* (NULL = the default stream).  Every entry point
 * returns an ed_status; ed_last_error() gives a thread-local message for the
 * last failure on the calling thread.
 *
 * The reference has no FFI: its boundary is header-only C++ templates over
 * thrust::device_vector.  Each declaration below cites the reference interface
 * it replaces; the headers under include/ed/ re-create those C++ classes on top of
 * this ABI.
 *
 * Layout conventions:
 *   basis:   ascending integers with popcount Q (colex order); address[] has 2^N
 *            entries, the rank inside the sector and 0 outside it.
 *   tables:  column-major ELL, slot k of row a at k*count + a, S = 10*Q + 1 slots
 *            (10*N + 1 for the full Fock space); complex tables interleave (re, im).
 *   vectors: psi[i_up*dim_dn + i_dn], FP64 or complex128 (interleaved re, im).
 */
#ifndef EDHIP_H
#define EDHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    ED_OK = 0,
    ED_EINVAL = 1,      /* bad argument */
    ED_ENOMEM = 2,      /* device or host allocation failed */
    ED_EBREAKDOWN = 3,  /* Lanczos breakdown that could not be handled */
    ED_ENOCONV = 4,     /* Lanczos did not meet its convergence test within max_steps */
    ED_EHIP = 5,        /* a HIP runtime call failed */
    ED_EUNSUPPORTED = 6,
    ED_ESLOT = 7        /* an ELL row ran out of slots (the reference writes past the table) */
} ed_status;

/* operator kinds for ed_tables_build, one per H_TBq method */
typedef enum {
    ED_OP_HOP = 0,       /* H_TBq::Make_H(i,j,t)        fermi_operator.h:100-115 */
    ED_OP_CURRENT = 1,   /* H_TBq::Make_J(i,j,t)        fermi_operator.h:117-138 */
    ED_OP_DENSITY = 2,   /* H_TBq::Make_N(i,v)          fermi_operator.h:140-149 */
    ED_OP_DENSITY2 = 3,  /* H_TBq::Insert_2Nop(i,j,v)   fermi_operator.h:216-225 */
    ED_OP_HOP_PHASE = 4, /* t e^{i phase} c+_i c_j + h.c. (complex tables; a Peierls twist --
                            no counterpart in the reference, whose tables are real) */
    ED_OP_FERMI1 = 5,    /* H_TBq::Insert_1Fop(dag_i, i, t)          fermi_operator.h:151-174:
                            t c^{dag_i}_i, sign (-1)^(occupied sites above i); full Fock space */
    ED_OP_FERMI2 = 6     /* H_TBq::Insert_2Fop(dag_i, i, dag_j, j, t) fermi_operator.h:176-214:
                            rows with occ_i, occ_j matching the two operators, sign
                            (-1)^(occupied sites strictly between i and j) */
} ed_op_kind;

typedef struct {
    int32_t kind; /* ed_op_kind */
    int32_t i, j; /* sites (j unused for ED_OP_DENSITY and ED_OP_FERMI1) */
    int32_t dag;  /* ED_OP_FERMI1/2 only: bit 0 = 1 if the operator on site i is an annihilator
                     ('-' in the reference, which acts on rows with site i occupied), 0 for a
                     creator ('+'); bit 1 the same for site j.  0 for the other kinds. */
    double value; /* t or v */
    double phase; /* ED_OP_HOP_PHASE only */
} ed_op;

const char *ed_last_error(void);
/* Device memory the library itself holds (its own allocations on all devices: operator tables,
 * layouts, solver work vectors, per-thread workspaces; not rocBLAS/rocSOLVER workspaces and not
 * the caller's buffers): current and peak bytes since load or the last reset.  reset_peak != 0
 * sets the peak to the current value after reading.  Host only. */
int ed_memory_stats(int64_t *current_bytes, int64_t *peak_bytes, int reset_peak);
int ed_version(void);
int ed_device_sync(void);

/* ---------------- basis + tables (fermi_operator.h) ---------------------------- */

/* count = C(N,Q) (2^N when Q < 0: the full-Fock H_TBq(N), fermi_operator.h:90-98) */
int ed_basis_count(int N, int Q, int64_t *count);

/* H_TBq::Labeling / Full_Labeling (fermi_operator.h:44-62, ctor 80-98):
 * d_state[count], d_address[2^N] (rank inside the sector, 0 elsewhere). */
int ed_basis_build(int N, int Q, int32_t *d_state, int32_t *d_address, void *stream);

/* LOW_MEM::H_TBq::Labeling (TotalSpin/include/fermi_operator_lowmem.h:50-64): the sector basis
 * of 64-bit states in the order std::prev_permutation visits them from the pattern with the Q
 * ones first (descending value of the bit-reversed state), for N <= 63.  d_state[count]
 * receives the states; the rank of a state is its index (the reference keeps a std::map).
 * Asynchronous. */
int ed_basis_build_lowmem(int N, int Q, uint64_t *d_state, void *stream);

/* H_TBq::Make_H / Make_J / Make_N / Insert_1Fop / Insert_2Fop / Insert_2Nop applied in call
 * order with the stacking rule (fermi_operator.h:64-225).  ops is a host array.  The
 * tables are cleared first (H_TBq::clear, fermi_operator.h:237-241), so one call
 * builds a whole operator; d_Ax holds count*S doubles (2x for complex tables).
 * Returns ED_ESLOT if a row overflows its S slots. */
int ed_tables_build(int N, int Q, const int32_t *d_state, const int32_t *d_address, int64_t count,
                    const ed_op *ops, int nops, int complex_tables, double *d_Ax, int32_t *d_Aj, void *stream);

/* ---------------- Hubbard operator (HubMatrix_GPU.cuh:30-182) ---------------- */

typedef struct ed_hubbard_s *ed_hubbard;

/* Hubbard(HqUP, HqDN, U [, dilution]) (HubMatrix_GPU.cuh:46-95).  Reads the ELL
 * tables (device) once and keeps a compact copy; the caller's tables may be freed
 * afterwards.  mask = dilution (-1: all sites).  Uvec: optional host array of N
 * per-site interactions that replaces U*popcount(up&dn&mask). */
int ed_hubbard_create(int N, int64_t dim_up, int S_up, const double *d_AxUP, const int32_t *d_AjUP,
                      const int32_t *d_stateUP, int64_t dim_dn, int S_dn, const double *d_AxDN,
                      const int32_t *d_AjDN, const int32_t *d_stateDN, int complex_tables, double U, int64_t mask,
                      const double *Uvec, ed_hubbard *out);
/* Hubbard(HqUP, HqDN, U [, dilution]) built straight from the two H_TBq call lists
 * (the Make_H / Make_J / Make_N / Insert_1Fop / Insert_2Fop / Insert_2Nop calls of
 * fermi_operator.h:100-225, in call order, as ed_tables_build takes them).  Builds the bases
 * and tables itself and, when each list has at most 64 calls and no Insert_1Fop / Insert_2Fop
 * call, applies H with hop descriptors generated on the fly from the call list (same matrix
 * elements, same per-element summation order as the tables); otherwise with the table
 * kernels.  Operators with Insert_1Fop / Insert_2Fop calls are not accepted by the batch
 * engine, the autotuner or the bilinear (gradient) entry points (ED_EUNSUPPORTED). */
int ed_hubbard_create_ops(int N, int Q_up, const ed_op *ops_up, int nops_up, int Q_dn, const ed_op *ops_dn,
                          int nops_dn, int complex_tables, double U, int64_t mask, const double *Uvec,
                          ed_hubbard *out);
int ed_hubbard_destroy(ed_hubbard h);
int ed_hubbard_info(ed_hubbard h, int64_t *dim_up, int64_t *dim_dn, int *complex_tables);

/* Hubbard::Hubbard_mv<M>(x, y) (HubMatrix_GPU.cuh:163-182): y = H x, y overwritten,
 * x != y.  complex_vectors selects M = complex<double>.  Asynchronous on stream. */
int ed_hubbard_apply(ed_hubbard h, int complex_vectors, const void *d_x, void *d_y, void *stream);

/* Tile shapes of the on-the-fly H.v (no counterpart in the reference, whose kernel has one
 * launch shape, HubMatrix_GPU.cuh:54-61).  slab_width: columns per pass-1 tile (32 to 256;
 * 16 B per lane when dim_dn is even); row_order: pass 1 visits the up rows in reverse
 * Cuthill-McKee order of the up hopping graph (built at creation for vectors of 1 GiB and more)
 * instead of ascending rank; up_table: pass 1 reads each row's hops from a table built at
 * creation (dim_up x calls words) instead of generating them per slab; dn_cached: pass 2 keeps
 * per-column call descriptors in LDS over dn_panels panels of dn_rows rows (else it regenerates
 * them per panel). */
typedef struct {
    int slab_width, dn_cached, dn_panels, dn_rows, row_order, up_table;
    double ms_per_apply;
} ed_hv_tuning;
/* Autotuner: times each candidate shape with `reps` applications of H to d_x (into d_y, which is
 * overwritten) on the stream and keeps the fastest in the handle; every candidate gives the same
 * y bit for bit.  Needs an operator from ed_hubbard_create_ops.  best (may be NULL) gets the
 * choice and its time. */
int ed_hubbard_autotune(ed_hubbard h, int complex_vectors, const void *d_x, void *d_y, int reps, ed_hv_tuning *best,
                        void *stream);

/* ---------------- dense full-spectrum path ---------------------------------------- */

/* Hubbard::initialize (HubMatrix_GPU.cuh:250-270): zero the n x n matrix d_A (complex128 when
 * complex_values).  Asynchronous. */
int ed_dense_zero(int64_t n, int complex_values, void *d_A, void *stream);
/* Hubbard::Write (HubMatrix_GPU.cuh:184-230, kernel write_): ADDS H into the dense D x D matrix
 * d_H, D = dim_up*dim_dn, row-major d_H[row*D + col] with row = i_up*dim_dn + i_dn, the
 * reference's indexing; per row the nonzero UP slots, then the nonzero DN slots, then the
 * interaction.  complex128 entries for an operator with complex tables.  D <= 65536.
 * Asynchronous. */
int ed_hubbard_write_dense(ed_hubbard h, void *d_H, void *stream);
/* H_TBq::Write (fermi_operator.h:227-235): d_M[j*count + Aj] += Ax over the nonzero slots of
 * the ELL tables of ed_tables_build (count <= 65536).  Asynchronous. */
int ed_tables_write_dense(int64_t count, int S, int complex_tables, const double *d_Ax, const int32_t *d_Aj,
                          void *d_M, void *stream);
/* cusolverDsyevd / cusolverZheevd (EIGEN_cusolver.h:16-63, 65-112), as tED<C>::Make_eigen and the
 * TotalSpinFull driver call them: all eigenpairs of the Hermitian n x n matrix d_A written by
 * the functions above (n <= 46340).  d_W[n] (device) receives the eigenvalues ascending and d_A
 * the unit eigenvectors, eigenvector k contiguous at d_A + k*n elements.  rocSOLVER syevd/heevd.
 * Synchronises; ED_ENOCONV when the solver reports a failure. */
int ed_dense_eigh(int64_t n, int complex_values, void *d_A, double *d_W, void *stream);
/* tED<C>::tED_time_evol (HubMatrix_GPU.cuh:784-797): d_phi += sum_k exp(-i W_k dt) <V_k|phi0> V_k
 * over the n eigenpairs from ed_dense_eigh (d_V real, or complex128 when complex_values); phi0
 * and phi are complex128 device vectors of n elements (phi accumulates, as in the reference:
 * zero it for exp(-iH dt) phi0).  Synchronises. */
int ed_ted_time_evol(int64_t n, int complex_values, const void *d_V, const double *d_W, const void *d_phi0,
                     double dt, void *d_phi, void *stream);

/* ---------------- slab layout (no counterpart in the reference) ------------------- */

/* The vector layout the Lanczos solver keeps internally for large real sectors: both species'
 * basis states renumbered in reverse Cuthill-McKee order of their hopping graphs (positions
 * pu, pd) and the state cut into slabs of 16 dn positions, amplitude (i_up, i_dn) at
 * (pd >> 4)*dim_up*16 + pu*W + (pd & 15) with W = 16 except in the last slab, which is only as
 * wide as the remaining dn positions (rounded up to even): *len = dim_up*dim_dn for even dim_dn
 * (else dim_up more);
 * padding entries zero.  H.v on it (ed_hubbard_apply_slab) gives the same values bit for bit as
 * ed_hubbard_apply (HubMatrix_GPU.cuh:114-182) on the reference layout, with 16 B gathers that
 * stay in L2.  enable = 1 builds the layout tables and a scratch vector (len doubles) in the
 * handle, after which ed_lanczos_eigen on real vectors runs its recurrence on this layout
 * (converting the start vector in and the Ritz vector out); enable = 0 frees them.  Needs a
 * real operator from ed_hubbard_create_ops with N <= 26.  ed_lanczos_eigen and
 * ed_lanczos_tridiagonal build it by themselves for real sectors of 2^27 amplitudes and more when
 * its scratch vector and every layout vector the solver allocates fit in free device memory with
 * 5 % of the card to spare (ed_slab_auto_decision; ED_SLAB_AUTO=0 turns that off).  enable = -1
 * only reports *len (0 when there is no layout). */
int ed_hubbard_slab_layout(ed_hubbard h, int enable, int64_t *len);
/* The solvers' decision above, host only: 1 when a dim_up x dim_dn real sector would get the slab
 * layout with free_bytes of total_bytes free and nvec more layout vectors for the solver
 * (eigen with a Ritz vector: 2 when dim_dn is even, the layout then has no padding and the
 * caller's f and evec carry layout vectors, else 4; eigenvalue only or tridiagonal: 3), else 0. */
int ed_slab_auto_decision(double free_bytes, double total_bytes, int64_t dim_up, int64_t dim_dn, int nvec);
/* reference layout psi[i_up*dim_dn + i_dn] -> slab layout (padding set to zero), and back */
int ed_slab_from_rowmajor(ed_hubbard h, const double *d_x, double *d_xl, void *stream);
int ed_slab_to_rowmajor(ed_hubbard h, const double *d_xl, double *d_x, void *stream);
/* y = H x on the slab layout (x != y, x padding zero; y padding is written zero).  Asynchronous. */
int ed_hubbard_apply_slab(ed_hubbard h, const double *d_xl, double *d_yl, void *stream);
/* Slab-layout autotuner (no counterpart in the reference; the reference has one launch shape,
 * HubMatrix_GPU.cuh:54-61): times the launch shapes of the slab H.v with `reps` applications of
 * H to d_xl (into d_yl, overwritten) and keeps the fastest in the layout -- first tile rows (32,
 * 64) x gathers in flight per lane (8, 16), then the tile order (plain, or runs of 2, 4, 8, 16
 * adjacent tiles per CU slot) with the best of those; every candidate gives the same y bit for
 * bit.  best (may be NULL) gets the choice and its time per H.v; candidate_ms holds the times in
 * the order tried ([32,8], [32,16], [64,8], [64,16] at order 8, then orders 0, 2, 4, 16). */
typedef struct {
    int tile_rows, gathers_in_flight, tile_order;
    double ms_per_apply;
    double candidate_ms[8];
} ed_slab_tuning;
int ed_hubbard_slab_autotune(ed_hubbard h, const double *d_xl, double *d_yl, int reps, ed_slab_tuning *best,
                             void *stream);
/* Sets the slab H.v launch shape directly (a choice the autotuner made earlier): tile rows 32 or
 * 64, gathers in flight 8 or 16, tile order 0, 2, 4, 8 or 16.  Host only; results do not depend on
 * it.  Default 32, 8, 8. */
int ed_hubbard_slab_set_shape(ed_hubbard h, int tile_rows, int gathers_in_flight, int tile_order);
/* Measurement probe (no counterpart in the reference): the line gathers of one pass of
 * ed_hubbard_apply_slab alone -- pass 0 the UP gathers over d_xl, pass 1 the DN gathers over the
 * layout's transposed scratch -- with the same launch shape (the autotuner's choice), grid, tile
 * order, LDS footprint and entry tables,
 * but without the value table, the FMAs and the stream loads and stores.  Its time is the cost
 * of that pass's gathers.  Asynchronous; writes nothing the caller sees. */
int ed_hubbard_slab_gather_ceiling(ed_hubbard h, int pass, const double *d_xl, void *stream);

/* ---------------- vector BLAS-1 (Thrust_Lin, HubMatrix_GPU.cuh:291-471) -------- */

/* GenRand + random_vector (HubMatrix_GPU.cuh:504-539): minstd_rand seeded with
 * `seed`, element idx = draw idx+1, U(-1,1); complex: two consecutive draws
 * (LanczosSolver.cuh:38-55).  normalize != 0 scales to unit norm. */
int ed_random_vector(int64_t n, uint32_t seed, int complex_vectors, void *d_v, int normalize, void *stream);
/* thrust_Dotprod (HubMatrix_GPU.cuh:451-471): out[0] = Re<a,b>, out[1] = Im<a,b> (0 for
 * real); deterministic reduction; synchronises the stream. */
int ed_dot(int64_t n, int complex_vectors, const void *d_a, const void *d_b, double *out, void *stream);
/* Allocation-free variant of ed_dot.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same kernels and reduction order as ed_dot, so the same value bit for bit, but the
 * block partials live in a fixed library-owned device array (16 MiB per GPU, held until the
 * process ends): n real elements (2n for complex vectors) must not exceed 2^32, else ED_EINVAL.
 * d_out is DEVICE memory for two doubles, {Re<a,b>, Im<a,b>} (Im written 0 for real vectors).
 * Asynchronous; allocates nothing and never synchronizes.  Calls that overlap on one GPU (other
 * streams or host threads) share the array and race: serialize them. */
int ed_dot_noalloc(int64_t n, int complex_vectors, const void *d_a, const void *d_b, double *d_out, void *stream);
/* y = alpha*x + y (real alpha) */
int ed_axpy(int64_t n, int complex_vectors, double alpha, const void *d_x, void *d_y, void *stream);
/* x = alpha*x */
int ed_scal(int64_t n, int complex_vectors, double alpha, void *d_x, void *stream);
/* normalize (HubMatrix_GPU.cuh:366-375); returns the squared norm before scaling */
int ed_normalize(int64_t n, int complex_vectors, void *d_x, double *norm2_out, void *stream);

/* ---------------- Lanczos (HubMatrix_GPU.cuh:477-613) ------------------------- */

typedef struct {
    int max_steps;   /* maxN (reference 1000) */
    double tol;      /* tolconv (reference 0.01*sqrt(eps)) */
    int check_every; /* convergence test cadence (reference 10) */
    int reorth;      /* 0: none (the reference); 1: second Gram-Schmidt pass of f against v_N
                      * (DGKS correction); 2: selective orthogonalisation (Parlett-Scott): the
                      * Lanczos basis is kept on the device, a Ritz analysis at every convergence
                      * check keeps the converged ("good") Ritz vectors, and every new f is
                      * orthogonalised against them (v_N too when one is found); the Ritz vector
                      * is then V z, with no second recurrence.  ED_ENOMEM when the basis does
                      * not fit. */
} ed_lanczos_cfg;

/* Lanczos<T>::eigen (HubMatrix_GPU.cuh:561-613).  d_f: start vector in (any norm),
 * clobbered.  d_evec: normalized Ritz vector out, rebuilt by re-running the
 * recurrence (NULL: eigenvalue only, no rebuild).  a_out/b_out (host, max_steps+2 entries, may be NULL) receive the
 * tridiagonal; *nsteps = N+1 steps kept.  Returns ED_ENOCONV (with the last Ritz
 * pair) when the test is not met by max_steps.  Synchronises the stream. */
int ed_lanczos_eigen(ed_hubbard h, int complex_vectors, const ed_lanczos_cfg *cfg, void *d_f, double *eval,
                     void *d_evec, double *a_out, double *b_out, int *nsteps, void *stream);

/* The nev lowest eigenvalues of H by one Lanczos run with selective orthogonalisation (the role
 * of cuSpectra::SymEigsSolver's "SmallestAlge" solve, TotalSpin/include/SymEigsSolver.cuh:332-359,
 * which restarts a 2 nev + 1 vector Krylov basis; here one recurrence runs on, its vectors kept
 * semi-orthogonal against the converged Ritz vectors up to the nev-th lowest as with reorth = 2).
 * Every check_every steps the nev lowest eigenvalues of T are taken with copies of converged
 * eigenvalues merged and converging ghost copies dropped (ed_tridiag_lanczos_lowest with
 * semi_orth = 1: the purges keep most copies from forming, not those of eigenvalues closer
 * together than the accuracy at which a Ritz vector counts as converged; a value the start
 * vector barely touches is kept), and each simple one must pass |b_{N+1} s_k[N]| < tol * anorm
 * (ed_lanczos_eigen's test); the lowest also needs |s_0[0]| > tol * anorm.  evals[nev]
 * ascending; d_evec (may be NULL) gets the normalised ground state: V s_0, refined by a plain
 * Lanczos run from it (ed_lanczos_eigen with the same test, Ritz vector rebuilt), since the
 * purges leave V s_0 off by about sqrt(eps); the refined value must agree with evals[0] to 1e-8
 * relative (ED_ENOCONV otherwise); *nsteps counts the steps of both runs.  cfg->reorth is ignored; nev 1..64.  The basis is kept on the device:
 * ED_ENOMEM when it does not fit.  Returns ED_ENOCONV (with the last Ritz values) when a test is
 * not met by max_steps.  Synchronises the stream. */
int ed_lanczos_lowest(ed_hubbard h, int complex_vectors, const ed_lanczos_cfg *cfg, void *d_f, int nev,
                      double *evals, void *d_evec, int *nsteps, void *stream);

/* Lanczos<T>::tridiagonal (HubMatrix_GPU.cuh:541-559): nsteps fixed steps from d_f
 * (not modified); a_out[nsteps], b_out[nsteps+1] with b_out[0] = |f|.  Real vectors run on
 * the slab layout when the operator has one, or gets one as in ed_lanczos_eigen. */
int ed_lanczos_tridiagonal(ed_hubbard h, int complex_vectors, int nsteps, const void *d_f, double *a_out,
                           double *b_out, void *stream);

/* Statistics of the calling thread's last ed_lanczos_eigen (no counterpart in the reference):
 * with reorth = 2, the good Ritz vectors kept, the Ritz analyses and the steps whose f was
 * orthogonalised against them; otherwise omega_max, the omega monitor's estimate of the largest
 * |v_j . v_k| of the run (ed_lanczos_orthogonality over its a and b).  Any pointer may be NULL. */
int ed_lanczos_so_stats(int *good, int *analyses, int *purges, double *omega_max);

/* Orthogonality monitor of a Lanczos run (no counterpart in the reference, whose recurrence
 * has no reorthogonalisation, HubMatrix_GPU.cuh:573-582): from a[0..n-1] and b[0..n-1] (b[0] the
 * start norm, b[k] coupling v_{k-1} and v_k) the omega recurrence (Simon 1984) estimates
 * omega_max[j] = max_{k<j} |v_j . v_k| for j = 0..n-1, assuming no reorthogonalisation.
 * Semi-orthogonality, omega <= sqrt(eps), keeps the tridiagonal's Ritz values accurate to
 * working precision.  Host only. */
int ed_lanczos_orthogonality(int n, const double *a, const double *b, double *omega_max);

/* lowest eigenpair of a symmetric tridiagonal (replaces LAPACKf77::dsted, MAGMA_LAPACK.h:296-326,
 * for the Lanczos test): d[n], e[n-1]; z[n] (may be NULL) gets the unit eigenvector. */
int ed_tridiag_lowest(int n, const double *d, const double *e, double *lambda, double *z);
/* all eigenvalues ascending */
int ed_tridiag_eigvals(int n, const double *d, const double *e, double *w);
/* all eigenpairs of a symmetric tridiagonal (LAPACKf77::dsted with compz = 'I', as
 * Krylov::Make_tridiagonal_eigen calls it, HubMatrix_GPU.cuh:690-699): d[n], e[n-1]; w[n]
 * ascending; z[n*n] column-major, z[j*n + i] = component i of the unit eigenvector of w[j]. */
int ed_tridiag_eigen(int n, const double *d, const double *e, double *w, double *z);
/* The nev lowest eigenvalues of a Lanczos tridiagonal (d[n], e[n-1]) with bres = b_{n+1}, the
 * coupling after the last step, by the filter ed_lanczos_lowest (semi_orth = 1: the recurrence
 * ran with selective orthogonalisation) and ed_batch_lowest_k (semi_orth = 0: without
 * reorthogonalisation) apply, after Cullum and Willoughby: copies within 1e-12 anorm merged,
 * ghosts dropped, values on eigenvalues of T without its first row and column (ghosts, or
 * eigenvalues the start vector barely touches) held back or dropped by the rules of each mode,
 * each kept simple value tested |bres s[n-1]| < tol * anorm.  out[nev] ascending (NaN past the values found); *found =
 * the number found; *conv = 1 when all nev were found and pass.  No counterpart in the reference
 * (its SymEigsSolver restarts instead of filtering); exported for tests of the filter. */
int ed_tridiag_lanczos_lowest(int n, const double *d, const double *e, double bres, int nev, double tol,
                              double anorm, int semi_orth, double *out, int *found, int *conv);

/* Krylov<C>::tridiagonal (HubMatrix_GPU.cuh:660-680): Lanczos recurrence from the complex128
 * vector d_phi0 (used as it is, not normalised, not modified) without reorthogonalisation, at
 * most max_steps steps, stopping after the first b_k < btol (the reference uses 1000, 1e-10).
 * a_out[max_steps], b_out[max_steps + 1] with b_out[0] = 0; *n_out = number of a's. */
int ed_krylov_tridiagonal(ed_hubbard h, const void *d_phi0, int max_steps, double btol, double *a_out,
                          double *b_out, int *n_out, void *stream);
/* Krylov<C>::Krylov_time_evol (HubMatrix_GPU.cuh:702-728): d_phi += sum_i c_i q_i with
 * c_i = sum_j exp(-i m_j dt) z[j*n] z[j*n+i], the Krylov vectors q_i of d_phi0 rebuilt by the
 * same recurrence from a[n], b[n+1] (from ed_krylov_tridiagonal) and m, z (from
 * ed_tridiag_eigen).  Complex128; d_phi accumulates (zero it for exp(-iH dt) phi0). */
int ed_krylov_time_evol(ed_hubbard h, const void *d_phi0, int n, const double *a, const double *b, const double *m,
                        const double *z, double dt, void *d_phi, void *stream);

/* ---------------- several lowest eigenpairs (cuSpectra) ------------------------- */

/* cuSpectra::SymEigsSolver<cuSpectra::Hubbard<T>>(op, nev, ncv) + compute(SmallestAlge, maxit,
 * tol) (MBL_test/include/SymEigsSolver.cuh:36-475, ArnoldiOp.cuh:132-219): the nev lowest
 * eigenvalues (ascending) by thick-restart Lanczos with full reorthogonalisation on a basis of
 * ncv device vectors (1 <= nev < ncv <= 64); maxit restarts at most; convergence as the
 * reference: |residual_i| <= tol * max(eps^(2/3), |theta_i|).  Start vector: the seed-`seed`
 * GenRand vector (the reference uses Spectra's SimpleRandom; eigenvalues do not depend on it).
 * d_evecs (NULL or nev vectors, vector i at offset i*D elements) receives the Ritz vectors.
 * Returns ED_ENOCONV when fewer than nev pairs converged (evals then hold the last Ritz values). */
int ed_eigs_lowest(ed_hubbard h, int complex_vectors, int nev, int ncv, int maxit, double tol, uint32_t seed,
                   double *evals, void *d_evecs, int *nconv, int *niter, void *stream);

/* cuSpectra::Hubbard<T>::perform_op(x_host, y_host) (HubMatrix_TotalSpin.cuh:181-322): y = H x
 * on host arrays of D elements; device buffers are allocated, filled and freed per call, as in
 * the reference (synchronous). */
int ed_hubbard_apply_host(ed_hubbard h, int complex_vectors, const void *x, void *y, void *stream);

/* ---------------- batched ground states ----------------------------------------- */

/* Many independent instances (sectors, twist fluxes, disorder realisations), each built by
 * ed_hubbard_create_ops, solved together: every Lanczos step of all active instances is one
 * launch per kernel over a shared tile list; after every ten steps (the convergence-check
 * cadence) a device kernel runs the convergence test of every active instance (the host test's
 * code, ed_tridiag.hpp) beside the next ten steps and switches converged instances off.  Fifty
 * steps with their tests are captured in one hipGraph (per buffer rotation phase) and replayed;
 * between replays the host reads only the number of instances still running.  Each instance
 * runs the recurrence and convergence test of Lanczos<T>::eigen (HubMatrix_GPU.cuh:561-613) from
 * the seed-`seed` start vector of its own length; only eigenvalues are produced.  No
 * counterpart in the reference, whose drivers loop over instances (e.g.
 * drude_v2_cal_Kagome_randomHop.cu:129-175). */
typedef struct ed_batch_s *ed_batch;
int ed_batch_create(int ninst, const ed_hubbard *ops, int complex_vectors, ed_batch *out);
int ed_batch_destroy(ed_batch b);
/* evals[ninst], nsteps[ninst], status[ninst] (ED_OK or ED_ENOCONV) out; use_graph = 0 launches
 * the same kernels without capture; 1 replays fifty-step graphs; 2 captures every step after the
 * first ten, up to cfg->max_steps, in one graph and launches it once (the steps after the last
 * instance has stopped find no tile to work on and return at once: set max_steps near the
 * expected step count; the library's own sweeps use 1, which is as fast).  The results do not
 * depend on it, bit for bit.  Synchronises. */
int ed_batch_lowest(ed_batch b, const ed_lanczos_cfg *cfg, uint32_t seed, int use_graph, double *evals,
                    int *nsteps, int *status, void *stream);
/* The nev lowest eigenvalues of every instance from one batched run (the SymEigsSolver
 * "SmallestAlge" role, SymEigsSolver.cuh:332-359, for many instances at once): every ten steps
 * a device test applies the filter below to each running instance (eigenvalues by Sturm
 * multisection) and switches off those whose nev lowest pass; an instance runs at most
 * cfg->max_steps steps (fewer when its recurrence ends in an invariant subspace).  Then its
 * tridiagonal is filtered on the host by the same test (ed_tridiag_lanczos_lowest with
 * semi_orth = 0: copies of converged eigenvalues merged; a simple value that is also an
 * eigenvalue of T without its first row and column -- a ghost, or an eigenvalue the start vector
 * barely touches, 1e-7 is enough -- holds the test back when it has converged, until its copy
 * forms, and when it lies more than 1e-3 anorm above the converged values below it, and is
 * dropped as a ghost in transit otherwise) and each of the nev lowest kept values must pass
 * ed_lanczos_eigen's residual test |b s_N| < tol * anorm.  evals[ninst * nev] (instance-major,
 * ascending, NaN past the values found), nsteps[ninst], status[ninst]: ED_OK, or ED_ENOCONV when
 * the test is not met.  d_evecs (NULL: none) holds ninst device vectors of the instances'
 * lengths (FP64 or complex128 as the batch): each instance with status ED_OK gets its normalised
 * ground state, the Ritz vector of T_n at the first check (every ten steps) where the lowest pair
 * passes the test -- where ed_lanczos_eigen would stop -- by a second batched run of the same
 * recurrence (the same a and b bit for bit) that accumulates sum z_k v_k, each instance for its
 * own n steps; an instance whose T_n value is not the filter's lowest (to 1e-8 relative) gets
 * ED_ENOCONV and no vector.  Synchronises. */
int ed_batch_lowest_k(ed_batch b, const ed_lanczos_cfg *cfg, int nev, uint32_t seed, int use_graph, double *evals,
                      int *nsteps, int *status, void *const *d_evecs, void *stream);
/* Panel shape of the batch kernels: DN rows per panel (4 or 8) x DN panels per workgroup (1-16),
 * UP rows per panel (1 or 4) x UP panels per workgroup (1-64); default 4 x 4, 1 x 4.  The
 * eigenvalues do not depend on it, bit for bit (same per-element sums; the dot partials are per
 * 4-row unit whatever the shape).  Host; rebuilds the tile tables. */
int ed_batch_set_shape(ed_batch b, int dn_rows, int dn_panels, int up_rows, int up_panels);
/* Autotuner (no counterpart in the reference): times a solve of `steps` (>= 10) recurrence steps of
 * every instance for each shape -- DN rows {4, 8} x DN panels {1, 2, 4, 8}, then UP rows {1, 4} x
 * UP panels {1, 2, 4, 8, 16} with the best DN shape -- and keeps the fastest.  best (may be NULL)
 * gets the choice, its time and every candidate's time in the order tried.  Synchronises. */
typedef struct {
    int dn_rows, dn_panels, up_rows, up_panels, steps;
    double seconds;
    double candidate_s[18];
} ed_batch_tuning;
int ed_batch_autotune(ed_batch b, int steps, ed_batch_tuning *best, void *stream);

/* ---------------- adjoint / Hellmann-Feynman --------------------------------------- */

/* For every call o of one species' list (species 0 = up, 1 = down) of an operator built by
 * ed_hubbard_create_ops: out[4o..4o+3] = (Re F_o, Im F_o, Re G_o, Im G_o) with
 *   F_o = <a| part of call o acting on rows with site i occupied |b>,
 *   G_o = <a| part acting on rows with site j occupied |b>
 * (matrix elements as built: sign * value * e^{+-i phase}).  H is linear in each call's value,
 * so dH/dt_o = (F_o + G_o)/t_o and the phase derivative is i (F_o - G_o).  Density calls
 * return F only.  These are the parameter cotangents of y = H(theta) x for a = grad_y, b = x,
 * and the Hellmann-Feynman derivatives for a = b = psi.  No counterpart in the reference.
 * Synchronises. */
int ed_bond_bilinear(ed_hubbard h, int species, int complex_vectors, const void *d_a, const void *d_b, double *out,
                     void *stream);
/* out[i] = Re sum over states with site i doubly occupied of conj(a) b, i < N (dH/dU_i). */
int ed_site_double_occupancy(ed_hubbard h, int complex_vectors, const void *d_a, const void *d_b, double *out,
                             void *stream);
/* out[sigma*N + i] = Re <a| n_{i sigma} |b> (sigma = 0 up, 1 down): the gradient of on-site
 * potentials eps_{i sigma} n_{i sigma} (the Make_N(i, eps) disorder terms of the MBL drivers,
 * Hubbard_EE_test.cu:119-140); no counterpart in the reference, which has no gradients. */
int ed_site_density(ed_hubbard h, int complex_vectors, const void *d_a, const void *d_b, double *out,
                    void *stream);

/* ---------------- observables -------------------------------------------------- */

/* TotalSpin::Hubbard (HubMatrix_TotalSpin.cuh:126-173): y = (S+S- + S-S+)/2 x, real or complex. */
typedef struct ed_s2_s *ed_s2;
int ed_s2_create(int N, int64_t dim_up, const int32_t *d_stateUP, const int32_t *d_addressUP, int64_t dim_dn,
                 const int32_t *d_stateDN, const int32_t *d_addressDN, ed_s2 *out);
int ed_s2_destroy(ed_s2 s);
int ed_s2_apply(ed_s2 s, int complex_vectors, const void *d_x, void *d_y, void *stream);

/* correlator_continued_fraction (HubMatrix_GPU.cuh:829-834), host: g[k] = norm / (w_k + E - a_0 - b_1^2/(...)) */
int ed_continued_fraction(int n, const double *a, const double *b, double energy, double norm, int nw,
                          const double *w_ri, double *g_ri);

/* Entanglement entropy of psi for the bipartition A = sites in maskA (EEmatHubAB_GPU +
 * cusolverDsvd/Zsvd + -sum s ln s, entangle_entropy_GPU.cuh:27-144, entangle_entropy.h:231-353,
 * Hubbard_EE_test.cu:77-83).  The (A,B) reshaped psi is block-diagonal in (N_A^up, N_A^dn);
 * each block rho_A = Psi Psi^T (Psi^H) is formed with FP64 MFMA and diagonalised
 * (values only).  out[0] = S_vN = -sum lambda ln lambda; out[1] = S_ref = -sum sigma ln sigma
 * with sigma from an SVD of each block (the reference's formula).  nsv_out (may be NULL):
 * number of nonzero Schmidt values.  Synchronises the stream.  When the largest block has 2048
 * rows or more (ED_EE_PARALLEL_MIN), the blocks are solved by four worker threads on streams of
 * their own, after the stream's earlier work; the results are those of the one-stream solve bit
 * for bit.  The same holds for the two calls below. */
int ed_entanglement_entropy(int N, int Q_up, int Q_dn, int64_t maskA, int complex_vectors, const void *d_psi,
                            double *out, int64_t *nsv_out, void *stream);
/* S_vN only (out[0] of ed_entanglement_entropy without the per-block SVDs, which dominate the
 * cost for blocks of a few hundred rows and more) */
int ed_entanglement_entropy_vn(int N, int Q_up, int Q_dn, int64_t maskA, int complex_vectors, const void *d_psi,
                               double *s_vn, void *stream);
/* S_vN of nstates states of the same sector and cut (no counterpart in the reference, whose drivers
 * take one state at a time): d_psis is a host array of nstates device pointers, s_vn (host)
 * gets one entropy per state.  Each block size of rho_A is solved once for a chunk of states (one
 * Gram launch, one strided-batched rocSOLVER eigenvalue call) instead of once per state; a
 * state's entropy is ed_entanglement_entropy_vn's sum over the same block eigenvalues (rocSOLVER's
 * batched routine may round them differently in the last bits).  Chunks hold at most half the
 * free device memory.  Synchronises the stream. */
int ed_entanglement_entropy_vn_batch(int N, int Q_up, int Q_dn, int64_t maskA, int complex_vectors, int nstates,
                                     const void *const *d_psis, double *s_vn, void *stream);
/* all eigenvalues of rho_A, descending, into lambda_out (capacity cap); *nlambda = total count */
int ed_entanglement_spectrum(int N, int Q_up, int Q_dn, int64_t maskA, int complex_vectors, const void *d_psi,
                             double *lambda_out, int64_t cap, int64_t *nlambda, void *stream);
/* All eigenvalues (ascending, into d_w[n]) of a real symmetric n x n matrix d_A (column-major, ld
 * lda; the lower triangle is read, or with upper != 0 the upper one, and the matrix is
 * overwritten) by a two-stage reduction: dense -> band (blocked Householder panels, rocSOLVER
 * geqrf/larft, BLAS-3 two-sided updates), band -> tridiagonal by bulge chasing, then Sturm
 * bisection of every eigenvalue.  The routine the entanglement calls use for rho_A blocks of 1024
 * rows and more in place of a one-stage dsyevd (the SVD role of entangle_entropy.h:231-353); no
 * counterpart in the reference.  Synchronises the stream. */
int ed_sym_eigvals(int64_t n, double *d_A, int64_t lda, int upper, double *d_w, void *stream);
/* Kernel-time statistics of the three calls above, summed over calls since the last reset (no
 * counterpart in the reference: a measurement aid).  *gram_ms: the rho_A Gram kernels' durations,
 * *gram_flops: their SYRK FLOPs (k (k+1) K per real k x K block, 4x that for complex), *eig_ms:
 * the durations of the per-block eigenvalue (and SVD) stage; each duration is timed by events on
 * the stream the work ran on.  Any pointer may be NULL; reset != 0 zeroes the sums after reading. */
int ed_entanglement_stats(double *gram_ms, double *gram_flops, double *eig_ms, int reset);

#ifdef __cplusplus
}
#endif

#endif /* EDHIP_H */

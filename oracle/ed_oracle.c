/*
 * ed_oracle.c -- CPU ORACLE for the exact-diagonalization hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the
 * reference arithmetic, which tests/ and __graft_entry__.smoke() check the
 * GPU results against and bench.py times as its cpu_baseline leg.  Nothing in
 * the product package links, loads or calls it; the product path fails
 * loudly when its HIP library is missing instead of falling back here.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root).  How each function is pinned (closed forms,
 * brute force, scipy/LAPACK fixtures, the reference's own fermi_operator.h
 * compiled into oracle/_ref) is listed in tests/test_oracle_*.py.
 *
 * Layout conventions (same as the reference):
 *   - single-species basis: ascending integers with popcount Q
 *     (fermi_operator.h:44-53); address[] has 2^N entries, 0 outside the sector.
 *   - operator tables: column-major ELL, slot k of row a at k*count+a,
 *     S = z*Q+1 slots with z = 10 (fermi_operator.h:11,86-87).
 *   - state vector: row-major psi[i_up*dim_dn + i_dn] (HubMatrix_GPU.cuh:141,149).
 *
 * Threads: the element loops of the H.v restatements use OpenMP over rows; every
 * reduction is done in fixed 4096-element chunks summed in index order, so results
 * do not depend on the thread count.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_Z 10 /* fermi_operator.h:11 */

enum { OR_OK = 0, OR_EINVAL = 1, OR_ENOMEM = 2, OR_EBREAKDOWN = 3, OR_ENOCONV = 4, OR_ESLOT = 7 };

/* operator kinds, one per H_TBq method (fermi_operator.h:100-225) */
enum {
    OR_OP_HOP = 0,       /* Make_H(i,j,t)           fermi_operator.h:100-115 */
    OR_OP_CURRENT = 1,   /* Make_J(i,j,t)           fermi_operator.h:117-138 */
    OR_OP_DENSITY = 2,   /* Make_N(i,v)             fermi_operator.h:140-149 */
    OR_OP_DENSITY2 = 3,  /* Insert_2Nop(i,j,v)      fermi_operator.h:216-225 */
    OR_OP_HOP_PHASE = 4, /* Make_H with a Peierls phase (complex tables only; no
                            counterpart in the reference, which has real tables) */
    OR_OP_FERMI1 = 5,    /* Insert_1Fop(dag,i,t)    fermi_operator.h:151-174 */
    OR_OP_FERMI2 = 6     /* Insert_2Fop(dag_i,i,dag_j,j,t) fermi_operator.h:176-214 */
};

/* Insert_1Fop / Insert_2Fop (fermi_operator.h:151-214).  dag bit 0 is i_bit ('+' -> 0: the
 * operator acts on rows with site i empty; '-' -> 1: occupied), bit 1 is j_bit.  Returns 1 and
 * the column and sign when row state n has an element.  Insert_1Fop's sign counts the occupied
 * sites strictly between i and N (count1_bt_sites(n, i_site, N), fermi_operator.h:166). */
static int fermi_entry(int N, int kind, int i, int j, int dag, int64_t n, const int32_t *address, int32_t *b,
                       double *sign);

/* ---- a1: count1, Kernighan popcount (fermi_operator.h:19-24) ---------------- */
int or_count1(int64_t state)
{
    int i;
    for (i = 0; state != 0; i++) state &= (state - 1);
    return i;
}

/* ---- a2: count1_bt_sites (fermi_operator.h:26-40): number of occupied sites
 * strictly between i and j, in either order. ------------------------------ */
int or_count1_bt(int64_t basis, int i_site, int j_site)
{
    int64_t b;
    if (j_site > i_site) {
        b = basis >> (i_site + 1);
        b = b % ((int64_t)1 << (j_site - i_site - 1));
    } else {
        b = basis >> (j_site + 1);
        b = b % ((int64_t)1 << (i_site - j_site - 1));
    }
    return or_count1(b);
}

static int fermi_entry(int N, int kind, int i, int j, int dag, int64_t n, const int32_t *address, int32_t *b,
                       double *sign)
{
    int i_bit = dag & 1, j_bit = (dag >> 1) & 1;
    int oi = (int)((n >> i) & 1);
    if (kind == OR_OP_FERMI1) {
        if (oi != i_bit) return 0;
        *sign = pow(-1.0, or_count1_bt(n, i, N));
        *b = address[n ^ ((int64_t)1 << i)];
        return 1;
    }
    int oj = (int)((n >> j) & 1);
    if (oi != i_bit || oj != j_bit) return 0;
    *sign = pow(-1.0, or_count1_bt(n, i, j));
    *b = address[n ^ (((int64_t)1 << i) | ((int64_t)1 << j))];
    return 1;
}

/* ---- a3: H_TBq::Labeling (fermi_operator.h:44-53) and Full_Labeling (55-62).
 * address must hold 2^N zero-initialised entries (thrust host_vector::resize
 * value-initialises, fermi_operator.h:83-84).  Q < 0 selects Full_Labeling.
 * Returns count. ------------------------------------------------------------- */
int64_t or_basis(int N, int Q, int32_t *state, int32_t *address)
{
    int64_t dim = (int64_t)1 << N, count = 0;
    for (int64_t n = 0; n < dim; n++) {
        if (Q < 0 || or_count1(n) == Q) {
            state[count] = (int32_t)n;
            address[n] = (int32_t)count;
            count++;
        }
    }
    return count;
}

/* ---- a4: H_TBq::stacking (fermi_operator.h:64-73): first slot whose value is
 * exactly 0.0 receives (value, column).  Overflow past the last slot is UB in
 * the reference (it walks off the array); here it returns OR_ESLOT. --------- */
static int stack_real(double *Ax, int32_t *Aj, int64_t count, int S, int64_t a, int32_t b, double sign,
                      double hopping)
{
    for (int k = 0; k < S; k++) {
        int64_t idx = (int64_t)k * count + a;
        if (Ax[idx] != 0.0) continue;
        Ax[idx] += sign * hopping;
        Aj[idx] = b;
        return OR_OK;
    }
    return OR_ESLOT;
}

static int stack_cplx(double complex *Ax, int32_t *Aj, int64_t count, int S, int64_t a, int32_t b, double sign,
                      double complex hopping)
{
    for (int k = 0; k < S; k++) {
        int64_t idx = (int64_t)k * count + a;
        if (Ax[idx] != 0.0) continue;
        Ax[idx] += sign * hopping;
        Aj[idx] = b;
        return OR_OK;
    }
    return OR_ESLOT;
}

/* ---- a5-a7: Make_H / Make_J / Make_N / Insert_2Nop applied in call order
 * (fermi_operator.h:100-149, 216-225).  Real tables.  Ax/Aj hold S*count
 * zero-initialised entries, S = 10*Q+1. --------------------------------------- */
int or_tables(int N, int Q, const int32_t *state, const int32_t *address, int64_t count, int nops,
              const int32_t *kind, const int32_t *si, const int32_t *sj, const double *val, double *Ax, int32_t *Aj,
              const int32_t *dag)
{
    int S = OR_Z * (Q < 0 ? N : Q) + 1;
    for (int o = 0; o < nops; o++) {
        int i = si[o], j = sj[o];
        double t = val[o];
        for (int64_t a = 0; a < count; a++) {
            int64_t n = state[a];
            int oi = (int)((n >> i) & 1), oj = (int)((n >> j) & 1);
            int st = OR_OK;
            if (kind[o] == OR_OP_HOP) {
                if (oi != oj) {
                    int fs = or_count1_bt(n, i, j);
                    int32_t b = address[n ^ (((int64_t)1 << i) | ((int64_t)1 << j))];
                    st = stack_real(Ax, Aj, count, S, a, b, pow(-1.0, fs), t);
                }
            } else if (kind[o] == OR_OP_CURRENT) {
                int fs = or_count1_bt(n, i, j);
                int32_t b = address[n ^ (((int64_t)1 << i) | ((int64_t)1 << j))];
                if (oi == 0 && oj == 1) st = stack_real(Ax, Aj, count, S, a, b, pow(-1.0, fs), t);
                else if (oi == 1 && oj == 0) st = stack_real(Ax, Aj, count, S, a, b, pow(-1.0, fs), -t);
            } else if (kind[o] == OR_OP_DENSITY) {
                if (oi == 1) st = stack_real(Ax, Aj, count, S, a, (int32_t)a, 1.0, t);
            } else if (kind[o] == OR_OP_DENSITY2) {
                if (oi == 1 && oj == 1) st = stack_real(Ax, Aj, count, S, a, (int32_t)a, 1.0, t);
            } else if (kind[o] == OR_OP_FERMI1 || kind[o] == OR_OP_FERMI2) {
                int32_t b;
                double sg;
                if (fermi_entry(N, kind[o], i, j, dag ? dag[o] : 0, n, address, &b, &sg))
                    st = stack_real(Ax, Aj, count, S, a, b, sg, t);
            } else {
                return OR_EINVAL;
            }
            if (st != OR_OK) return st;
        }
    }
    return OR_OK;
}

/* Complex tables: the same slot rules with complex values.  OR_OP_HOP_PHASE(i,j,t,phi)
 * is t*e^{i phi} c^+_i c_j + t*e^{-i phi} c^+_j c_i: a row with site i occupied and j
 * empty gets sign*t*e^{+i phi}; the opposite row gets sign*t*e^{-i phi}.  val_im is
 * the phase phi for OR_OP_HOP_PHASE and ignored otherwise.  Ax is interleaved (re,im). */
int or_tables_z(int N, int Q, const int32_t *state, const int32_t *address, int64_t count, int nops,
                const int32_t *kind, const int32_t *si, const int32_t *sj, const double *val, const double *phase,
                double *Ax_ri, int32_t *Aj, const int32_t *dag)
{
    int S = OR_Z * (Q < 0 ? N : Q) + 1;
    double complex *Ax = (double complex *)Ax_ri;
    for (int o = 0; o < nops; o++) {
        int i = si[o], j = sj[o];
        double t = val[o];
        for (int64_t a = 0; a < count; a++) {
            int64_t n = state[a];
            int oi = (int)((n >> i) & 1), oj = (int)((n >> j) & 1);
            int st = OR_OK;
            int32_t b = 0;
            if (kind[o] == OR_OP_HOP || kind[o] == OR_OP_CURRENT || kind[o] == OR_OP_HOP_PHASE)
                b = address[n ^ (((int64_t)1 << i) | ((int64_t)1 << j))];
            if (kind[o] == OR_OP_HOP) {
                if (oi != oj) st = stack_cplx(Ax, Aj, count, S, a, b, pow(-1.0, or_count1_bt(n, i, j)), t);
            } else if (kind[o] == OR_OP_HOP_PHASE) {
                double ph = phase[o];
                if (oi == 1 && oj == 0)
                    st = stack_cplx(Ax, Aj, count, S, a, b, pow(-1.0, or_count1_bt(n, i, j)), t * cexp(I * ph));
                else if (oi == 0 && oj == 1)
                    st = stack_cplx(Ax, Aj, count, S, a, b, pow(-1.0, or_count1_bt(n, i, j)), t * cexp(-I * ph));
            } else if (kind[o] == OR_OP_CURRENT) {
                int fs = or_count1_bt(n, i, j);
                if (oi == 0 && oj == 1) st = stack_cplx(Ax, Aj, count, S, a, b, pow(-1.0, fs), t);
                else if (oi == 1 && oj == 0) st = stack_cplx(Ax, Aj, count, S, a, b, pow(-1.0, fs), -t);
            } else if (kind[o] == OR_OP_DENSITY) {
                if (oi == 1) st = stack_cplx(Ax, Aj, count, S, a, (int32_t)a, 1.0, t);
            } else if (kind[o] == OR_OP_DENSITY2) {
                if (oi == 1 && oj == 1) st = stack_cplx(Ax, Aj, count, S, a, (int32_t)a, 1.0, t);
            } else if (kind[o] == OR_OP_FERMI1 || kind[o] == OR_OP_FERMI2) {
                double sg;
                if (fermi_entry(N, kind[o], i, j, dag ? dag[o] : 0, n, address, &b, &sg))
                    st = stack_cplx(Ax, Aj, count, S, a, b, sg, t);
            } else {
                return OR_EINVAL;
            }
            if (st != OR_OK) return st;
        }
    }
    return OR_OK;
}

/* ---- a11: Hubbard_mv_ (HubMatrix_GPU.cuh:114-161), real.
 * Per element: UP slots 0..S_up-1 skipping Ax==0 (140-142), then DN slots
 * 0..S_dn-1 including zeros (146-150), then the U term (152-156).  Uvec != NULL
 * replaces U*popcount(up&dn&mask) by sum_i Uvec[i] over doubly occupied sites
 * (ascending i) -- the per-site interaction used by the disorder/gradient
 * workloads, an extension of the reference's uniform U with its dilution mask. */
static double diag_u(int64_t up, int64_t dn, double U, int64_t mask, const double *Uvec)
{
    int64_t dbl = up & dn & mask;
    if (!Uvec) return U * (double)or_count1(dbl);
    double s = 0.0;
    for (int i = 0; dbl; i++, dbl >>= 1)
        if (dbl & 1) s += Uvec[i];
    return s;
}

/* rows [row0, row0+nrows) of y = H x (y indexed like x); used whole by or_hubbard_apply and
 * on a bounded row sample for the CPU throughput baseline */
void or_hubbard_apply_rows(int64_t dim_up, int64_t dim_dn, int S_up, const double *AxUP, const int32_t *AjUP,
                           int S_dn, const double *AxDN, const int32_t *AjDN, const int32_t *stateUP,
                           const int32_t *stateDN, double U, int64_t mask, const double *Uvec, const double *x,
                           double *y, int64_t row0, int64_t nrows)
{
    const int64_t row_end = (row0 + nrows < dim_up) ? row0 + nrows : dim_up;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t u = row0; u < row_end; u++) {
        for (int64_t d = 0; d < dim_dn; d++) {
            double sum = 0.0;
            for (int s = 0; s < S_up; s++) {
                double a = AxUP[(int64_t)s * dim_up + u];
                if (a != 0) sum += a * x[(int64_t)AjUP[(int64_t)s * dim_up + u] * dim_dn + d];
            }
            for (int s = 0; s < S_dn; s++) {
                double a = AxDN[(int64_t)s * dim_dn + d];
                int64_t col = AjDN[(int64_t)s * dim_dn + d];
                sum += a * x[u * dim_dn + col];
            }
            if (U != 0 || Uvec) sum += diag_u(stateUP[u], stateDN[d], U, mask, Uvec) * x[u * dim_dn + d];
            y[u * dim_dn + d] = sum;
        }
    }
}

void or_hubbard_apply(int64_t dim_up, int64_t dim_dn, int S_up, const double *AxUP, const int32_t *AjUP, int S_dn,
                      const double *AxDN, const int32_t *AjDN, const int32_t *stateUP, const int32_t *stateDN,
                      double U, int64_t mask, const double *Uvec, const double *x, double *y)
{
    or_hubbard_apply_rows(dim_up, dim_dn, S_up, AxUP, AjUP, S_dn, AxDN, AjDN, stateUP, stateDN, U, mask, Uvec, x, y,
                          0, dim_up);
}

/* Complex H.v with complex tables (the Thrust complex<double> instantiation of
 * Hubbard_mv_, HubMatrix_GPU.cuh:114-161, extended to complex matrix elements). */
void or_hubbard_apply_z(int64_t dim_up, int64_t dim_dn, int S_up, const double *AxUP_ri, const int32_t *AjUP,
                        int S_dn, const double *AxDN_ri, const int32_t *AjDN, const int32_t *stateUP,
                        const int32_t *stateDN, double U, int64_t mask, const double *Uvec, const double *x_ri,
                        double *y_ri)
{
    const double complex *AxUP = (const double complex *)AxUP_ri, *AxDN = (const double complex *)AxDN_ri;
    const double complex *x = (const double complex *)x_ri;
    double complex *y = (double complex *)y_ri;
#pragma omp parallel for schedule(static)
    for (int64_t u = 0; u < dim_up; u++) {
        for (int64_t d = 0; d < dim_dn; d++) {
            double complex sum = 0.0;
            for (int s = 0; s < S_up; s++) {
                double complex a = AxUP[(int64_t)s * dim_up + u];
                if (a != 0) sum += a * x[(int64_t)AjUP[(int64_t)s * dim_up + u] * dim_dn + d];
            }
            for (int s = 0; s < S_dn; s++) {
                double complex a = AxDN[(int64_t)s * dim_dn + d];
                int64_t col = AjDN[(int64_t)s * dim_dn + d];
                sum += a * x[u * dim_dn + col];
            }
            if (U != 0 || Uvec) sum += diag_u(stateUP[u], stateDN[d], U, mask, Uvec) * x[u * dim_dn + d];
            y[u * dim_dn + d] = sum;
        }
    }
}

/* ---- a18: TotalSpin kernel::Hubbard_mv_ (HubMatrix_TotalSpin.cuh:20-66):
 * y = (S+S- + S-S+)/2 x by the gather form with the exchanged-state sign. ---- */
void or_s2_apply(int N, int64_t dim_up, int64_t dim_dn, const int32_t *stateUP, const int32_t *stateDN,
                 const int32_t *addressUP, const int32_t *addressDN, const double *x, double *y)
{
#pragma omp parallel for schedule(static)
    for (int64_t sv = 0; sv < dim_up; sv++) {
        for (int64_t id = 0; id < dim_dn; id++) {
            double sum = 0.0;
            int64_t up = stateUP[sv], dn = stateDN[id];
            for (int i = 0; i < N; ++i) {
                int iup = (int)((up >> i) & 1), idn = (int)((dn >> i) & 1);
                if (iup != idn) {
                    for (int j = i + 1; j < N; ++j) {
                        int jup = (int)((up >> j) & 1), jdn = (int)((dn >> j) & 1);
                        int64_t ex = ((int64_t)1 << i) | ((int64_t)1 << j);
                        if (iup == jdn && idn == jup) {
                            int64_t up_ = up ^ ex, dn_ = dn ^ ex;
                            int fs = or_count1_bt(up_, i, j) + or_count1_bt(dn_, i, j);
                            double xv = x[(int64_t)addressUP[up_] * dim_dn + addressDN[dn_]];
                            if (fs % 2 == 0) sum -= xv;
                            else sum += xv;
                        }
                    }
                }
            }
            double coef = 0.5 * or_count1(up ^ dn);
            sum += coef * x[sv * dim_dn + id];
            y[sv * dim_dn + id] = sum;
        }
    }
}

/* ---- a13: GenRand + random_vector (HubMatrix_GPU.cuh:504-539).
 * thrust::default_random_engine is minstd_rand (a = 48271, c = 0, m = 2^31-1),
 * seeded with 235; discard(idx) then one draw gives x_{idx+1} = 235*48271^(idx+1)
 * mod m; uniform_real_distribution(-1,1) maps it to ((x-1)/(m-1))*2 - 1.
 * The complex start (LanczosSolver.cuh:38-55) takes two consecutive draws.
 * No normalisation here: or_normalize does that. --------------------------- */
static uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) { return (a * b) % m; }

static uint64_t powmod(uint64_t base, uint64_t e, uint64_t m)
{
    uint64_t r = 1;
    base %= m;
    while (e) {
        if (e & 1) r = mulmod(r, base, m);
        base = mulmod(base, base, m);
        e >>= 1;
    }
    return r;
}

static double lcg_uniform(uint64_t x)
{
    double r = (double)(x - 1);
    r /= 2147483646.0;
    return r * 2.0 + (-1.0);
}

void or_random_vector(int64_t n, uint32_t seed, double *v)
{
    const uint64_t m = 2147483647ULL, a = 48271ULL;
    uint64_t s = seed % m;
    if (s == 0) s = 1;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) v[i] = lcg_uniform(mulmod(s, powmod(a, (uint64_t)i + 1, m), m));
}

void or_random_vector_z(int64_t n, uint32_t seed, double *v_ri)
{
    const uint64_t m = 2147483647ULL, a = 48271ULL;
    uint64_t s = seed % m;
    if (s == 0) s = 1;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        uint64_t x1 = mulmod(s, powmod(a, (uint64_t)i + 1, m), m);
        uint64_t x2 = mulmod(x1, a, m);
        v_ri[2 * i] = lcg_uniform(x1);
        v_ri[2 * i + 1] = lcg_uniform(x2);
    }
}

/* ---- a12: Thrust_Lin (HubMatrix_GPU.cuh:366-471, thrust_util.cuh) --------- */
#define CHUNK 4096
static double dot_r(int64_t n, const double *a, const double *b)
{
    int64_t nc = (n + CHUNK - 1) / CHUNK;
    double *part = (double *)malloc(sizeof(double) * (size_t)(nc ? nc : 1));
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nc; c++) {
        double s = 0.0;
        int64_t e = (c + 1) * CHUNK < n ? (c + 1) * CHUNK : n;
        for (int64_t i = c * CHUNK; i < e; i++) s += a[i] * b[i];
        part[c] = s;
    }
    double s = 0.0;
    for (int64_t c = 0; c < nc; c++) s += part[c];
    free(part);
    return s;
}

/* complex: returns <a|b> = sum conj(a) b as (re, im) */
static double complex dot_z(int64_t n, const double complex *a, const double complex *b)
{
    int64_t nc = (n + CHUNK - 1) / CHUNK;
    double complex *part = (double complex *)malloc(sizeof(double complex) * (size_t)(nc ? nc : 1));
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nc; c++) {
        double complex s = 0.0;
        int64_t e = (c + 1) * CHUNK < n ? (c + 1) * CHUNK : n;
        for (int64_t i = c * CHUNK; i < e; i++) s += conj(a[i]) * b[i];
        part[c] = s;
    }
    double complex s = 0.0;
    for (int64_t c = 0; c < nc; c++) s += part[c];
    free(part);
    return s;
}

double or_dot(int64_t n, const double *a, const double *b) { return dot_r(n, a, b); }

void or_dot_z(int64_t n, const double *a_ri, const double *b_ri, double *out_ri)
{
    double complex s = dot_z(n, (const double complex *)a_ri, (const double complex *)b_ri);
    out_ri[0] = creal(s);
    out_ri[1] = cimag(s);
}

/* normalize (HubMatrix_GPU.cuh:366-375): v *= 1/sqrt(sum v^2) */
void or_normalize(int64_t n, double *v)
{
    double s = dot_r(n, v, v), inv = 1. / sqrt(s);
    for (int64_t i = 0; i < n; i++) v[i] = inv * v[i];
}

void or_normalize_z(int64_t n, double *v_ri)
{
    double s = creal(dot_z(n, (const double complex *)v_ri, (const double complex *)v_ri)), inv = 1. / sqrt(s);
    for (int64_t i = 0; i < 2 * n; i++) v_ri[i] = inv * v_ri[i];
}

/* ---- a15: replacement for LAPACKf77::dsted = dstedc('I') (MAGMA_LAPACK.h:296-326).
 * The oracle needs the lowest eigenpair of the symmetric tridiagonal T (diag d,
 * off-diagonal e).  Eigenvalue: Sturm-sequence bisection (Golub & Van Loan,
 * Matrix Computations, sec. 8.4.1) to full precision.  Eigenvector: inverse
 * iteration with Gaussian elimination with partial pivoting on T - lambda I
 * (Golub & Van Loan sec. 8.2.2), two sweeps, normalised to unit 2-norm.  The sign
 * is fixed so the largest-magnitude component is positive (dstedc's sign is
 * arbitrary; only |z[0]|, |z[n-1]| and the Ritz vector up to sign are used). --- */
static int sturm_count(int n, const double *d, const double *e, double x)
{
    /* number of eigenvalues < x */
    int c = 0;
    double q = d[0] - x;
    if (q < 0) c++;
    for (int i = 1; i < n; i++) {
        double qq = q;
        if (fabs(qq) < 1e-300) qq = (qq < 0 ? -1e-300 : 1e-300);
        q = d[i] - x - e[i - 1] * e[i - 1] / qq;
        if (q < 0) c++;
    }
    return c;
}

double or_tridiag_lowest(int n, const double *d, const double *e, double *z)
{
    double lo = d[0], hi = d[0];
    for (int i = 0; i < n; i++) {
        double r = (i > 0 ? fabs(e[i - 1]) : 0) + (i < n - 1 ? fabs(e[i]) : 0);
        if (d[i] - r < lo) lo = d[i] - r;
        if (d[i] + r > hi) hi = d[i] + r;
    }
    double span = hi - lo;
    lo -= 1e-12 * (span + 1);
    hi += 1e-12 * (span + 1);
    for (int it = 0; it < 200; it++) {
        double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (sturm_count(n, d, e, mid) >= 1) hi = mid;
        else lo = mid;
    }
    double lam = 0.5 * (lo + hi);
    if (!z) return lam;
    if (n == 1) {
        z[0] = 1.0;
        return lam;
    }
    /* inverse iteration: (T - lam I) x_{k+1} = x_k, solved by Gaussian elimination
     * with partial pivoting on the tridiagonal (the dgttrf/dgttrs elimination order) */
    double *dl = malloc(sizeof(double) * n), *dg = malloc(sizeof(double) * n), *du = malloc(sizeof(double) * n);
    double *du2 = malloc(sizeof(double) * n), *x = malloc(sizeof(double) * n);
    int *piv = malloc(sizeof(int) * n);
    double tiny = 1e-15 * (fabs(lam) + span + 1e-300);
    for (int i = 0; i < n; i++) {
        dg[i] = d[i] - lam;
        dl[i] = (i < n - 1) ? e[i] : 0.0;
        du[i] = (i < n - 1) ? e[i] : 0.0;
        du2[i] = 0.0;
        piv[i] = 0;
    }
    for (int i = 0; i < n - 1; i++) {
        if (fabs(dg[i]) >= fabs(dl[i])) {
            if (dg[i] == 0.0) dg[i] = tiny;
            double fact = dl[i] / dg[i];
            dl[i] = fact;
            dg[i + 1] -= fact * du[i];
        } else {
            double fact = dg[i] / dl[i];
            dg[i] = dl[i];
            dl[i] = fact;
            double tmp = du[i];
            du[i] = dg[i + 1];
            dg[i + 1] = tmp - fact * dg[i + 1];
            if (i < n - 2) {
                du2[i] = du[i + 1];
                du[i + 1] = -fact * du[i + 1];
            }
            piv[i] = 1;
        }
    }
    if (dg[n - 1] == 0.0) dg[n - 1] = tiny;
    for (int i = 0; i < n; i++) x[i] = 1.0;
    for (int sweep = 0; sweep < 3; sweep++) {
        for (int i = 0; i < n - 1; i++) {
            if (!piv[i]) x[i + 1] -= dl[i] * x[i];
            else {
                double t = x[i];
                x[i] = x[i + 1];
                x[i + 1] = t - dl[i] * x[i];
            }
        }
        x[n - 1] /= dg[n - 1];
        x[n - 2] = (x[n - 2] - du[n - 2] * x[n - 1]) / dg[n - 2];
        for (int i = n - 3; i >= 0; i--) x[i] = (x[i] - du[i] * x[i + 1] - du2[i] * x[i + 2]) / dg[i];
        double nrm = 0.0;
        for (int i = 0; i < n; i++) nrm += x[i] * x[i];
        nrm = sqrt(nrm);
        for (int i = 0; i < n; i++) x[i] /= nrm;
    }
    int imax = 0;
    for (int i = 1; i < n; i++)
        if (fabs(x[i]) > fabs(x[imax])) imax = i;
    double sg = x[imax] < 0 ? -1.0 : 1.0;
    for (int i = 0; i < n; i++) z[i] = sg * x[i];
    free(dl);
    free(dg);
    free(du);
    free(du2);
    free(x);
    free(piv);
    return lam;
}

/* all eigenvalues ascending, by bisection for each index (used by tests) */
void or_tridiag_eigvals(int n, const double *d, const double *e, double *w)
{
    double lo = d[0], hi = d[0];
    for (int i = 0; i < n; i++) {
        double r = (i > 0 ? fabs(e[i - 1]) : 0) + (i < n - 1 ? fabs(e[i]) : 0);
        if (d[i] - r < lo) lo = d[i] - r;
        if (d[i] + r > hi) hi = d[i] + r;
    }
    double span = hi - lo;
    lo -= 1e-12 * (span + 1);
    hi += 1e-12 * (span + 1);
    for (int k = 0; k < n; k++) {
        double l = lo, h = hi;
        for (int it = 0; it < 200; it++) {
            double mid = 0.5 * (l + h);
            if (mid <= l || mid >= h) break;
            if (sturm_count(n, d, e, mid) >= k + 1) h = mid;
            else l = mid;
        }
        w[k] = 0.5 * (l + h);
    }
}

/* ---- a14: Lanczos<T>::eigen (HubMatrix_GPU.cuh:561-613), real and complex.
 * An opaque operator context drives or_hubbard_apply(_z).  Buffers:
 * f (start vector in, overwritten), evec (out).  a[], b[] must hold maxN+1
 * entries.  Returns status; *nsteps = N+1 steps kept for the Ritz rebuild. */
typedef struct {
    int64_t dim_up, dim_dn;
    int S_up, S_dn;
    const double *AxUP, *AxDN; /* real or interleaved complex */
    const int32_t *AjUP, *AjDN, *stateUP, *stateDN;
    double U;
    int64_t mask;
    const double *Uvec;
    int cplx;
} or_op;

static void op_apply(const or_op *op, const double *x, double *y)
{
    if (op->cplx)
        or_hubbard_apply_z(op->dim_up, op->dim_dn, op->S_up, op->AxUP, op->AjUP, op->S_dn, op->AxDN, op->AjDN,
                           op->stateUP, op->stateDN, op->U, op->mask, op->Uvec, x, y);
    else
        or_hubbard_apply(op->dim_up, op->dim_dn, op->S_up, op->AxUP, op->AjUP, op->S_dn, op->AxDN, op->AjDN,
                         op->stateUP, op->stateDN, op->U, op->mask, op->Uvec, x, y);
}

/* one recurrence step on vectors of nr = n (real) or 2n (complex interleaved) doubles:
 * v1 = f*(1/b); f = H v1; f -= b v0 (N != 0); a = Re<v1,f>; f -= a v1; returns sum |f|^2 */
static double lanczos_step(const or_op *op, int64_t n, int N, double bN, double *f, double *v0, double *v1,
                           double *a_out)
{
    int64_t nr = op->cplx ? 2 * n : n;
    double inv = 1.0 / bN;
    for (int64_t i = 0; i < nr; i++) v1[i] = inv * f[i];
    op_apply(op, v1, f);
    if (N != 0)
        for (int64_t i = 0; i < nr; i++) f[i] = -bN * v0[i] + f[i];
    double a = dot_r(nr, v1, f); /* Re<v1,f> for interleaved complex equals the real dot of the pairs */
    *a_out = a;
    for (int64_t i = 0; i < nr; i++) f[i] = -a * v1[i] + f[i];
    memcpy(v0, v1, sizeof(double) * (size_t)nr);
    return dot_r(nr, f, f);
}

int or_lanczos_eigen(const or_op *op, int maxN, double tolconv, double *f, double *eval, double *evec, double *a,
                     double *b, int *nsteps)
{
    int64_t n = op->dim_up * op->dim_dn, nr = op->cplx ? 2 * n : n;
    double *v0 = calloc((size_t)nr, sizeof(double)), *v1 = calloc((size_t)nr, sizeof(double));
    double *z = NULL, anorm = 0.0;
    int N, conv = 0, status = OR_OK, lastN = -1;
    if (!v0 || !v1) return OR_ENOMEM;
    memcpy(evec, f, sizeof(double) * (size_t)nr); /* evec = f (563) keeps the start vector */
    b[0] = sqrt(dot_r(nr, f, f));
    z = malloc(sizeof(double) * (size_t)(maxN + 1));
    for (N = 0; N < maxN; ++N) {
        double nf2 = lanczos_step(op, n, N, b[N], f, v0, v1, &a[N]);
        b[N + 1] = sqrt(nf2);
        if (N == 0) anorm = fabs(a[0] + b[1]);
        else anorm = fmax(anorm, b[N] + fabs(a[N] + b[N + 1]));
        if (N % 10 == 9 || N == maxN - 1 || b[N + 1] == 0.0) {
            *eval = or_tridiag_lowest(N + 1, a, b + 1, z);
            lastN = N;
            if (b[N + 1] == 0.0) { /* invariant subspace: T is exact (the reference divides by 0) */
                conv = 1;
                break;
            }
            if (fabs(b[N + 1] * z[N]) < tolconv * anorm && fabs(z[0]) > tolconv * anorm) {
                conv = 1;
                break;
            }
        }
    }
    if (!conv) status = OR_ENOCONV;
    N = lastN;
    /* Ritz vector by re-running the recurrence from the saved start (600-612) */
    memcpy(f, evec, sizeof(double) * (size_t)nr);
    memset(evec, 0, sizeof(double) * (size_t)nr);
    memset(v0, 0, sizeof(double) * (size_t)nr);
    for (int i = 0; i < N + 1; ++i) {
        double inv = 1.0 / b[i];
        for (int64_t k = 0; k < nr; k++) v1[k] = inv * f[k];
        op_apply(op, v1, f);
        if (i != 0)
            for (int64_t k = 0; k < nr; k++) f[k] = -b[i] * v0[k] + f[k];
        for (int64_t k = 0; k < nr; k++) f[k] = -a[i] * v1[k] + f[k];
        memcpy(v0, v1, sizeof(double) * (size_t)nr);
        for (int64_t k = 0; k < nr; k++) evec[k] = z[i] * v1[k] + evec[k];
    }
    double s = dot_r(nr, evec, evec), inv = 1. / sqrt(s);
    for (int64_t k = 0; k < nr; k++) evec[k] = inv * evec[k];
    *nsteps = N + 1;
    free(v0);
    free(v1);
    free(z);
    return status;
}

/* ---- a14, two-vector form for sectors whose three D-vectors do not fit host memory
 * (18-site Kagome: D = 2.36e9, 18.9 GB per vector).  Same recurrence, start vector
 * (seed-235 GenRand, normalized) and convergence test as or_lanczos_eigen, eigenvalue
 * only.  The buffer holding v0 receives f = H v1 - b v0 row by row (each row of f needs
 * only v1 and the same row of v0), so two D-vectors suffice. */
/* compact copy of one species' ELL table: the nonzero slots of each row, in slot order */
typedef struct {
    int64_t *start;
    double *val;
    int32_t *col;
} csr_t;

static int csr_of(int64_t count, int S, const double *Ax, const int32_t *Aj, csr_t *c)
{
    int64_t nnz = 0;
    for (int64_t r = 0; r < count; r++)
        for (int s = 0; s < S; s++) nnz += Ax[(int64_t)s * count + r] != 0;
    c->start = malloc(sizeof(int64_t) * (size_t)(count + 1));
    c->val = malloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
    c->col = malloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
    if (!c->start || !c->val || !c->col) return OR_ENOMEM;
    int64_t k = 0;
    for (int64_t r = 0; r < count; r++) {
        c->start[r] = k;
        for (int s = 0; s < S; s++) {
            double a = Ax[(int64_t)s * count + r];
            if (a != 0) {
                c->val[k] = a;
                c->col[k++] = Aj[(int64_t)s * count + r];
            }
        }
    }
    c->start[count] = k;
    return OR_OK;
}

static void csr_free(csr_t *c)
{
    free(c->start);
    free(c->val);
    free(c->col);
}

/* y = H x - c y, row by row.  Per element the same summation order as or_hubbard_apply_rows:
 * UP slots in order (zeros skipped, as there), DN slots in order (the zero slots that
 * or_hubbard_apply_rows adds contribute exact zeros and are skipped here), then U.  The UP
 * sums run with the element index innermost (a row of x at a time), which keeps each
 * element's order. */
static void apply_minus_rows(const or_op *op, const csr_t *cu, const csr_t *cd, const double *x, double *y, double c)
{
    const int64_t du = op->dim_up, dd = op->dim_dn;
#pragma omp parallel
    {
        double *t = malloc(sizeof(double) * (size_t)dd);
#pragma omp for schedule(dynamic, 1)
        for (int64_t u = 0; u < du; u++) {
            for (int64_t d = 0; d < dd; d++) t[d] = 0.0;
            for (int64_t k = cu->start[u]; k < cu->start[u + 1]; k++) {
                const double a = cu->val[k];
                const double *xr = x + (int64_t)cu->col[k] * dd;
                for (int64_t d = 0; d < dd; d++) t[d] += a * xr[d];
            }
            const double *xu = x + u * dd;
            for (int64_t d = 0; d < dd; d++) {
                double sum = t[d];
                for (int64_t k = cd->start[d]; k < cd->start[d + 1]; k++) sum += cd->val[k] * xu[cd->col[k]];
                if (op->U != 0 || op->Uvec) sum += diag_u(op->stateUP[u], op->stateDN[d], op->U, op->mask, op->Uvec) * xu[d];
                y[u * dd + d] = (c != 0.0) ? -c * y[u * dd + d] + sum : sum;
            }
        }
        free(t);
    }
}

int or_lanczos_lowest_2vec(const or_op *op, int maxN, double tolconv, uint32_t seed, double *eval, double *a,
                           double *b, int *nsteps, int verbose)
{
    int64_t n = op->dim_up * op->dim_dn;
    if (op->cplx) return OR_EINVAL;
    double *P = malloc(sizeof(double) * (size_t)n), *V = malloc(sizeof(double) * (size_t)n);
    double *z = malloc(sizeof(double) * (size_t)(maxN + 1)), anorm = 0.0;
    int N, status = OR_ENOCONV;
    csr_t cu, cd;
    if (!P || !V || !z || csr_of(op->dim_up, op->S_up, op->AxUP, op->AjUP, &cu) ||
        csr_of(op->dim_dn, op->S_dn, op->AxDN, op->AjDN, &cd)) {
        free(P);
        free(V);
        free(z);
        return OR_ENOMEM;
    }
    or_random_vector(n, seed, V);
    or_normalize(n, V); /* random_vector (HubMatrix_GPU.cuh:531-539) */
    b[0] = sqrt(dot_r(n, V, V));
    *nsteps = 0;
    for (N = 0; N < maxN; ++N) {
        double inv = 1.0 / b[N];
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) V[i] = inv * V[i]; /* v1 = f * (1/b) */
        apply_minus_rows(op, &cu, &cd, V, P, N != 0 ? b[N] : 0.0); /* P: v0 -> f = H v1 - b v0 */
        a[N] = dot_r(n, V, P);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) P[i] = -a[N] * V[i] + P[i];
        b[N + 1] = sqrt(dot_r(n, P, P));
        double *t = P; /* v0 <- v1, f stays in the other buffer */
        P = V;
        V = t;
        if (N == 0) anorm = fabs(a[0] + b[1]);
        else anorm = fmax(anorm, b[N] + fabs(a[N] + b[N + 1]));
        if (N % 10 == 9 || N == maxN - 1 || b[N + 1] == 0.0) {
            *eval = or_tridiag_lowest(N + 1, a, b + 1, z);
            *nsteps = N + 1;
            if (verbose) {
                fprintf(stderr, "[oracle 2vec] step %d E0 %.15f |b z_N| %.3e\n", N + 1, *eval, fabs(b[N + 1] * z[N]));
                fflush(stderr);
            }
            if (b[N + 1] == 0.0 ||
                (fabs(b[N + 1] * z[N]) < tolconv * anorm && fabs(z[0]) > tolconv * anorm)) {
                status = OR_OK;
                break;
            }
        }
    }
    free(P);
    free(V);
    free(z);
    csr_free(&cu);
    csr_free(&cd);
    return status;
}

/* ---- a16: Lanczos<T>::tridiagonal (HubMatrix_GPU.cuh:541-559): maxN fixed steps,
 * a[0..maxN-1], b[0..maxN] (b[0] = |f|).  f is not modified (passed by value). */
int or_lanczos_tridiagonal(const or_op *op, int maxN, const double *f_in, double *a, double *b)
{
    int64_t n = op->dim_up * op->dim_dn, nr = op->cplx ? 2 * n : n;
    double *f = malloc(sizeof(double) * (size_t)nr), *v0 = calloc((size_t)nr, sizeof(double)),
           *v1 = calloc((size_t)nr, sizeof(double));
    if (!f || !v0 || !v1) return OR_ENOMEM;
    memcpy(f, f_in, sizeof(double) * (size_t)nr);
    b[0] = sqrt(dot_r(nr, f, f));
    for (int N = 0; N < maxN; ++N) b[N + 1] = sqrt(lanczos_step(op, n, N, b[N], f, v0, v1, &a[N]));
    free(f);
    free(v0);
    free(v1);
    return OR_OK;
}

/* ---- a16: correlator_continued_fraction loop (HubMatrix_GPU.cuh:829-834):
 * g(w) = norm / (w+E-a0 - b1^2/(w+E-a1 - ...)), evaluated from the bottom. */
void or_continued_fraction(int n, const double *a, const double *b, double energy, double norm, int nw,
                           const double *w_ri, double *g_ri)
{
    for (int idx = 0; idx < nw; idx++) {
        double complex w = w_ri[2 * idx] + I * w_ri[2 * idx + 1];
        double complex gw = 1.0 / (w + energy - a[n - 1]);
        for (int i = n - 2; i >= 0; i--) gw = 1.0 / (w + energy - a[i] - gw * b[i + 1] * b[i + 1]);
        double complex g = norm * gw;
        g_ri[2 * idx] = creal(g);
        g_ri[2 * idx + 1] = cimag(g);
    }
}

/* ---- a17: drude_weight_Jx Kubo combination (drude_v2_cal_Kagome_randomHop.cu:183-220)
 * given Lambda(omega_k) and <K>: sigma_k = -Im Lambda_k / (pi omega_k);
 * Vertex = sum_{omega_k > cutoff} sigma_k * step; D = (-K - 2 Vertex)/VOL. */
double or_drude_combine(int nw, const double *w_ri, const double *lambda_ri, double kinetic, double omega_step,
                        double omega_cutoff, double vol)
{
    double vertex = 0.0;
    for (int i = 0; i < nw; i++) {
        double wr = w_ri[2 * i];
        double sigma = -1.0 / (M_PI * wr) * lambda_ri[2 * i + 1];
        if (wr > omega_cutoff) vertex += sigma * omega_step;
    }
    return (-kinetic - 2.0 * vertex) / vol;
}

/* ---- convenience entry for ctypes: build an operator context ---------------- */
or_op *or_op_new(int64_t dim_up, int64_t dim_dn, int S_up, const double *AxUP, const int32_t *AjUP, int S_dn,
                 const double *AxDN, const int32_t *AjDN, const int32_t *stateUP, const int32_t *stateDN, double U,
                 int64_t mask, const double *Uvec, int cplx)
{
    or_op *op = calloc(1, sizeof(or_op));
    op->dim_up = dim_up;
    op->dim_dn = dim_dn;
    op->S_up = S_up;
    op->S_dn = S_dn;
    op->AxUP = AxUP;
    op->AjUP = AjUP;
    op->AxDN = AxDN;
    op->AjDN = AjDN;
    op->stateUP = stateUP;
    op->stateDN = stateDN;
    op->U = U;
    op->mask = mask;
    op->Uvec = Uvec;
    op->cplx = cplx;
    return op;
}

void or_op_free(or_op *op) { free(op); }

void or_op_apply(const or_op *op, const double *x, double *y) { op_apply(op, x, y); }

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

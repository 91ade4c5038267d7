// ed_hv_otf_body.hpp -- device bodies of the on-the-fly H.v passes, shared by the
// single-operator kernels (ed_hv_otf.hip) and the batched multi-instance kernels
// (ed_batch.hip).  See ed_hv_otf.hip for the algorithm.
#pragma once

#include "ed_common.hpp"
#include "ed_hubbard.hpp"

namespace ed {

template <typename V>
struct VO;
template <>
struct VO<double> {
    static __device__ __forceinline__ double zero() { return 0.0; }
    static __device__ __forceinline__ double scale(double s, double x) { return s * x; }
    static __device__ __forceinline__ double fmr(double a, double x, double acc) { return fma(a, x, acc); }
    static __device__ __forceinline__ double dotre(double a, double b) { return a * b; }
};
template <>
struct VO<cplx> {
    static __device__ __forceinline__ cplx zero() { return cplx{0.0, 0.0}; }
    static __device__ __forceinline__ cplx scale(double s, cplx x) { return cplx{s * x.re, s * x.im}; }
    static __device__ __forceinline__ cplx fmr(double a, cplx x, cplx acc)
    {
        return cplx{fma(a, x.re, acc.re), fma(a, x.im, acc.im)};
    }
    static __device__ __forceinline__ double dotre(cplx a, cplx b) { return a.re * b.re + a.im * b.im; }
};

__device__ __forceinline__ double omac(double a, double x, double acc) { return fma(a, x, acc); }
__device__ __forceinline__ cplx omac(double a, cplx x, cplx acc) { return cplx{fma(a, x.re, acc.re), fma(a, x.im, acc.im)}; }
__device__ __forceinline__ cplx omac(cplx a, cplx x, cplx acc)
{
    return cplx{fma(a.re, x.re, fma(-a.im, x.im, acc.re)), fma(a.re, x.im, fma(a.im, x.re, acc.im))};
}

template <typename TV>
__device__ __forceinline__ TV tv_make(double re, double im);
template <>
__device__ __forceinline__ double tv_make<double>(double re, double) { return re; }
template <>
__device__ __forceinline__ cplx tv_make<cplx>(double re, double im) { return cplx{re, im}; }

template <typename TV>
__device__ __forceinline__ TV tv_neg(TV a);
template <>
__device__ __forceinline__ double tv_neg<double>(double a) { return -a; }
template <>
__device__ __forceinline__ cplx tv_neg<cplx>(cplx a) { return cplx{-a.re, -a.im}; }

// Contribution of call o to the row of state s (row index `self`): returns true and
// (col, value) when the call has a matrix element in this row.  Values are the table
// values of fermi_operator.h: sign * t for Make_H, sign * (+-t) for Make_J, t for Make_N /
// Insert_2Nop, sign * t e^{+-i phi} for the phase hop.
template <typename TV>
__device__ __forceinline__ bool op_entry(const OtfOp &op, uint32_t s, int32_t self, const int32_t *__restrict__ addr,
                                         int32_t &col, TV &val)
{
    const uint32_t oi = (s >> op.i) & 1u, oj = (s >> op.j) & 1u;
    switch (op.kind) {
    case ED_OP_DENSITY:
        if (!oi) return false;
        col = self;
        val = tv_make<TV>(op.vf_re, op.vf_im);
        return true;
    case ED_OP_DENSITY2:
        if (!(oi & oj)) return false;
        col = self;
        val = tv_make<TV>(op.vf_re, op.vf_im);
        return true;
    default: {
        if (oi == oj) return false;
        col = addr[s ^ op.m];
        const TV v = oi ? tv_make<TV>(op.vf_re, op.vf_im) : tv_make<TV>(op.vb_re, op.vb_im);
        val = (__popc(s & op.between) & 1) ? tv_neg<TV>(v) : v;
        return true;
    }
    }
}


// ---------------------------------------------------------------------------------
// pass 1 body: UP term of logical tile t (slab-major).  A tile is 256/TPR rows x TPR*VEC
// columns: TPR threads per row, each owning VEC consecutive columns (VEC = 2 loads 16 B per
// lane for FP64).  Tile row r is up row up_order[r0 + r] (or r0 + r without an order).  The
// (column, value) list of each row comes from the precomputed entry table when there is one
// (a coalesced copy into LDS); else rows are evaluated one per wave (rows r = wave, wave + 4,
// ...): the lanes take one operator call each, a ballot orders the hopping ones into LDS, and
// each wave loads the states of all its rows before it looks up any target address.
// ---------------------------------------------------------------------------------
template <typename V, int VEC>
struct VecLoad;
template <typename V>
struct VecLoad<V, 1> {
    static __device__ __forceinline__ void load(const V *p, V *o) { o[0] = *p; }
    static __device__ __forceinline__ void store(V *p, const V *o) { *p = o[0]; }
};
template <>
struct VecLoad<double, 2> {
    static __device__ __forceinline__ void load(const double *p, double *o)
    {
        const double2 v = *reinterpret_cast<const double2 *>(p);
        o[0] = v.x;
        o[1] = v.y;
    }
    static __device__ __forceinline__ void store(double *p, const double *o)
    {
        *reinterpret_cast<double2 *>(p) = make_double2(o[0], o[1]);
    }
};
template <>
struct VecLoad<cplx, 2> {
    static __device__ __forceinline__ void load(const cplx *p, cplx *o)
    {
        o[0] = p[0];
        o[1] = p[1];
    }
    static __device__ __forceinline__ void store(cplx *p, const cplx *o)
    {
        p[0] = o[0];
        p[1] = o[1];
    }
};

template <typename V, typename TV, bool SCALE, int TPR, int VEC = 1>
__device__ __forceinline__ void otf_up_body(const OtfArgs &p, int64_t t, int64_t nrp, const V *__restrict__ x, V *out,
                                            double sc, TV *s_val, int32_t *s_col, int *s_cnt, int32_t *s_row,
                                            TV *s_tab)
{
    constexpr int ORB = 256 / TPR, OW = TPR * VEC, RPW = (ORB + 3) / 4;
    const int nops = p.up.nops;
    const int64_t slab = t / nrp, rp = t - slab * nrp;
    const int64_t u0 = rp * ORB;
    if (p.up_tab) {
        // precomputed entries: up row u holds up_cnt[u] words (column << 8 | value code) at
        // up_tab[u * nops ...], value code 4 o + k -> k = 0: vf, 1: -vf, 2: vb, 3: -vb
        for (int q = threadIdx.x; q < 4 * nops; q += 256) {
            const OtfOp op = p.up.ops[q >> 2];
            const int k = q & 3;
            const TV v = (k < 2) ? tv_make<TV>(op.vf_re, op.vf_im) : tv_make<TV>(op.vb_re, op.vb_im);
            s_tab[q] = (k & 1) ? tv_neg<TV>(v) : v;
        }
        if (threadIdx.x < ORB) {
            const int64_t q = u0 + threadIdx.x;
            const int32_t row = q < p.dim_up ? (p.up_order ? p.up_order[q] : (int32_t)q) : -1;
            s_row[threadIdx.x] = row;
            s_cnt[threadIdx.x] = row >= 0 ? p.up_cnt[row] : 0;
        }
        __syncthreads();
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        for (int r = wave; r < ORB; r += 4) {
            const int32_t row = s_row[r];
            if (row >= 0 && lane < s_cnt[r]) {
                const uint32_t e = p.up_tab[(int64_t)row * nops + lane];
                s_col[r * nops + lane] = (int32_t)(e >> 8);
                s_val[r * nops + lane] = s_tab[e & 0xffu];
            }
        }
    } else {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        OtfOp op{};
        if (lane < nops) op = p.up.ops[lane];
        int32_t ur[RPW];
        uint32_t st[RPW];
#pragma unroll
        for (int k = 0; k < RPW; ++k) {
            const int64_t q = u0 + wave + 4 * k;
            ur[k] = -1;
            if (wave + 4 * k < ORB && q < p.dim_up) ur[k] = p.up_order ? p.up_order[q] : (int32_t)q;
        }
#pragma unroll
        for (int k = 0; k < RPW; ++k) st[k] = ur[k] >= 0 ? (uint32_t)p.state_up[ur[k]] : 0u;
#pragma unroll
        for (int k = 0; k < RPW; ++k) {
            const int r = wave + 4 * k;
            bool has = false;
            int32_t col = 0;
            TV val{};
            if (ur[k] >= 0 && lane < nops) has = op_entry<TV>(op, st[k], ur[k], p.up.addr, col, val);
            const uint64_t bal = __ballot(has);
            if (has) {
                const int pos = __popcll(bal & ((1ull << lane) - 1ull));
                s_col[r * nops + pos] = col;
                s_val[r * nops + pos] = val;
            }
            if (lane == 0 && r < ORB) {
                s_cnt[r] = __popcll(bal);
                s_row[r] = ur[k];
            }
        }
    }
    __syncthreads();
    const int r = threadIdx.x / TPR, c = (threadIdx.x % TPR) * VEC;
    const int64_t u = s_row[r], d = slab * OW + c;
    if (u < 0 || d >= p.dim_dn) return;
    const int n = s_cnt[r];
    const int32_t *cl = s_col + r * nops;
    const TV *vl = s_val + r * nops;
    if (VEC == 1 || d + VEC > p.dim_dn) { // scalar columns (also the ragged last slab)
        for (int q = 0; q < VEC && d + q < p.dim_dn; ++q) {
            V acc = VO<V>::zero();
            for (int j0 = 0; j0 < n; j0 += 8) {
                V xv[8];
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (j0 + k < n) xv[k] = x[(int64_t)cl[j0 + k] * p.dim_dn + d + q];
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (j0 + k < n) {
                        V xk = xv[k];
                        if (SCALE) xk = VO<V>::scale(sc, xk);
                        acc = omac(vl[j0 + k], xk, acc);
                    }
            }
            out[u * p.dim_dn + d + q] = acc;
        }
        return;
    }
    // VEC consecutive columns per lane, loaded together (the launcher only picks VEC = 2 when
    // dim_dn is even, so every row chunk is 16 B aligned); the FMAs stay in call order
    V acc[VEC];
#pragma unroll
    for (int q = 0; q < VEC; ++q) acc[q] = VO<V>::zero();
    for (int j0 = 0; j0 < n; j0 += 8) {
        V xv[8][VEC];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (j0 + k < n) VecLoad<V, VEC>::load(x + (int64_t)cl[j0 + k] * p.dim_dn + d, xv[k]);
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (j0 + k < n) {
#pragma unroll
                for (int q = 0; q < VEC; ++q) {
                    V xk = xv[k][q];
                    if (SCALE) xk = VO<V>::scale(sc, xk);
                    acc[q] = omac(vl[j0 + k], xk, acc[q]);
                }
            }
    }
    VecLoad<V, VEC>::store(out + u * p.dim_dn + d, acc);
}

__device__ __forceinline__ double diag_of(int32_t su, int32_t sd, const OtfArgs &p)
{
    const uint32_t dbl = (uint32_t)(su & sd) & (uint32_t)p.mask;
    if (!p.Uvec) return p.U * (double)__popc(dbl);
    double s = 0.0;
    uint32_t m = dbl;
    while (m) {
        const int i = __ffs(m) - 1;
        s += p.Uvec[i];
        m &= m - 1;
    }
    return s;
}

// ---------------------------------------------------------------------------------
// pass 2 body: DN term + diagonal + epilogue of logical tile t (R rows x 256 columns);
// returns the block's Re<v,y> partial (valid in thread 0) when HV_DOT is set
// ---------------------------------------------------------------------------------
template <typename V, typename TV, int R, int FLAGS, bool MASKED = false>
__device__ __forceinline__ double otf_dn_body(const OtfArgs &p, int64_t t, int64_t ncb, bool has_up,
                                              const V *__restrict__ x, const V *ypart, V *y, V *vout, const V *w,
                                              double sc, double c, double *s_red)
{
    const int64_t rp = t / ncb, cb = t - rp * ncb;
    const int64_t u0 = rp * R, d = cb * 256 + threadIdx.x;
    const int64_t dim_dn = p.dim_dn;
    double dot = 0.0;
    if (d < dim_dn) {
        V acc[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int64_t u = u0 + r;
            acc[r] = (has_up && u < p.dim_up) ? ypart[u * dim_dn + d] : VO<V>::zero();
        }
        const uint32_t sd = (uint32_t)p.state_dn[d];
        const int nops = p.dn.nops;
        // calls in batches of 4: the 4 rank lookups, then 4*R gathers, are issued before the
        // FMAs, which stay in call order; a call without an element in this column points at
        // the own column with value 0 (adds an exact 0, as the reference's empty DN slots do)
        constexpr int KB = 4;
        for (int o0 = 0; o0 < nops; o0 += KB) {
            int32_t col[KB];
            TV a[KB];
            bool act[KB];
#pragma unroll
            for (int k = 0; k < KB; ++k) {
                col[k] = (int32_t)d;
                a[k] = tv_make<TV>(0.0, 0.0);
                act[k] = false;
                if (o0 + k < nops) {
                    int32_t ck;
                    TV ak;
                    if (op_entry<TV>(p.dn.ops[o0 + k], sd, (int32_t)d, p.dn.addr, ck, ak)) {
                        col[k] = ck;
                        a[k] = ak;
                        act[k] = true;
                    }
                }
            }
            V xv[KB][R];
            if (MASKED) {
                // only calls with an element in this column load; a call no lane of the wave
                // needs issues no load at all (the exact zeros of the reference's empty slots
                // are not added)
#pragma unroll
                for (int k = 0; k < KB; ++k) {
                    if (act[k]) {
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            const int64_t u = (u0 + r < p.dim_up) ? u0 + r : u0;
                            xv[k][r] = x[u * dim_dn + col[k]];
                        }
                    }
                }
#pragma unroll
                for (int k = 0; k < KB; ++k) {
                    if (act[k]) {
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            V xk = xv[k][r];
                            if (FLAGS & HV_SCALE) xk = VO<V>::scale(sc, xk);
                            acc[r] = omac(a[k], xk, acc[r]);
                        }
                    }
                }
                continue;
            }
#pragma unroll
            for (int k = 0; k < KB; ++k)
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int64_t u = (u0 + r < p.dim_up) ? u0 + r : u0;
                    xv[k][r] = x[u * dim_dn + col[k]];
                }
#pragma unroll
            for (int k = 0; k < KB; ++k)
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    V xk = xv[k][r];
                    if (FLAGS & HV_SCALE) xk = VO<V>::scale(sc, xk);
                    acc[r] = omac(a[k], xk, acc[r]);
                }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int64_t u = u0 + r;
            if (u >= p.dim_up) continue;
            const int64_t i = u * dim_dn + d;
            V xd = x[i];
            if (FLAGS & HV_SCALE) xd = VO<V>::scale(sc, xd);
            if (p.has_diag) acc[r] = VO<V>::fmr(diag_of(p.state_up[u], (int32_t)sd, p), xd, acc[r]);
            V o = acc[r];
            if (FLAGS & HV_AXPY) o = VO<V>::fmr(c, w[i], o);
            y[i] = o;
            if (FLAGS & HV_WRITEV) vout[i] = xd;
            if (FLAGS & HV_DOT) dot += VO<V>::dotre(xd, o);
        }
    }
    if (FLAGS & HV_DOT) return block_sum<256>(dot, s_red);
    return 0.0;
}

} // namespace ed

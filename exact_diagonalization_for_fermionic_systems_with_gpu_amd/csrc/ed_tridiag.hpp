// ed_tridiag.hpp -- the lowest eigenpair of a symmetric tridiagonal T (diagonal d[n], off-diagonal
// e[n-1]) by Sturm bisection and inverse iteration, for the host (ed_lanczos_eigen, the dsted role
// of MAGMA_LAPACK.h:296-326) and for the device (the batch engine's convergence test,
// ed_batch.hip).  One source for both, compiled without floating-point contraction, so that the
// two give the same bits from the same inputs (each body turns contraction off).
#pragma once

#include <cmath>
#include <cstdint>
#include <limits>

#include <hip/hip_runtime.h>

namespace ed {

// number of eigenvalues below x (Sturm sequence of the LDL^T pivots)
__host__ __device__ inline int tri_sturm_below(int n, const double *d, const double *e, double x)
{
#pragma clang fp contract(off)
    int c = 0;
    double q = d[0] - x;
    if (q < 0) ++c;
    for (int i = 1; i < n; ++i) {
        if (q == 0.0) q = std::numeric_limits<double>::min();
        q = (d[i] - x) - e[i - 1] * e[i - 1] / q;
        if (q < 0) ++c;
    }
    return c;
}

// the k-th smallest eigenvalue by bisection of the Gershgorin interval down to adjacent doubles
__host__ __device__ inline double tri_bisect_kth(int n, const double *d, const double *e, int k)
{
#pragma clang fp contract(off)
    double lo = d[0], hi = d[0];
    for (int i = 0; i < n; ++i) {
        const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i < n - 1 ? std::fabs(e[i]) : 0.0);
        lo = std::fmin(lo, d[i] - r);
        hi = std::fmax(hi, d[i] + r);
    }
    const double pad = 1e-12 * (hi - lo + 1.0);
    lo -= pad;
    hi += pad;
    for (int it = 0; it < 256; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (!(mid > lo && mid < hi)) break;
        if (tri_sturm_below(n, d, e, mid) > k) hi = mid;
        else lo = mid;
    }
    return 0.5 * (lo + hi);
}

// inverse iteration on T - lam I, elimination with partial pivoting, three sweeps; z normalised,
// sign such that the largest component is positive.  ws: 5 n doubles, sw: n flags.
__host__ __device__ inline void tri_inverse_iteration(int n, const double *d, const double *e, double lam, double *z,
                                                      double *ws, char *sw)
{
#pragma clang fp contract(off)
    if (n == 1) {
        z[0] = 1.0;
        return;
    }
    double *lo = ws, *dg = ws + n, *up = ws + 2 * n, *up2 = ws + 3 * n, *x = ws + 4 * n;
    double scale = std::fabs(lam);
    for (int i = 0; i < n; ++i) scale = std::fmax(scale, std::fabs(d[i]));
    const double tiny = std::numeric_limits<double>::epsilon() * (scale + 1e-300);
    for (int i = 0; i < n; ++i) {
        dg[i] = d[i] - lam;
        lo[i] = i < n - 1 ? e[i] : 0.0;
        up[i] = i < n - 1 ? e[i] : 0.0;
        up2[i] = 0.0;
        x[i] = 1.0;
        sw[i] = 0;
    }
    for (int i = 0; i < n - 1; ++i) {
        if (std::fabs(dg[i]) >= std::fabs(lo[i])) {
            if (dg[i] == 0.0) dg[i] = tiny;
            const double m = lo[i] / dg[i];
            lo[i] = m;
            dg[i + 1] -= m * up[i];
        } else {
            const double m = dg[i] / lo[i];
            dg[i] = lo[i];
            lo[i] = m;
            const double t = up[i];
            up[i] = dg[i + 1];
            dg[i + 1] = t - m * dg[i + 1];
            if (i < n - 2) {
                up2[i] = up[i + 1];
                up[i + 1] = -m * up[i + 1];
            }
            sw[i] = 1;
        }
    }
    if (dg[n - 1] == 0.0) dg[n - 1] = tiny;
    for (int sweep = 0; sweep < 3; ++sweep) {
        for (int i = 0; i < n - 1; ++i) {
            if (!sw[i]) {
                x[i + 1] -= lo[i] * x[i];
            } else {
                const double t = x[i];
                x[i] = x[i + 1];
                x[i + 1] = t - lo[i] * x[i];
            }
        }
        x[n - 1] /= dg[n - 1];
        x[n - 2] = (x[n - 2] - up[n - 2] * x[n - 1]) / dg[n - 2];
        for (int i = n - 3; i >= 0; --i) x[i] = (x[i] - up[i] * x[i + 1] - up2[i] * x[i + 2]) / dg[i];
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += x[i] * x[i];
        s = 1.0 / std::sqrt(s);
        for (int i = 0; i < n; ++i) x[i] *= s;
    }
    int imax = 0;
    for (int i = 1; i < n; ++i)
        if (std::fabs(x[i]) > std::fabs(x[imax])) imax = i;
    const double sg = x[imax] < 0 ? -1.0 : 1.0;
    for (int i = 0; i < n; ++i) z[i] = sg * x[i];
}

} // namespace ed

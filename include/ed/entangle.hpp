// ed/entangle.hpp -- EEmatHubAB_GPU<T> (MBL_test/include/entangle_entropy_GPU.cuh:27-144) with
// the cusolverDsvd/Zsvd + entropy step of the drivers (entangle_entropy.h:231-353,
// Hubbard_EE_test.cu:77-83) folded in.  The reference writes the whole rowD x colD matrix
// M[(a_up, a_dn), (b_up, b_dn)] = psi and takes its SVD; that matrix is block-diagonal in the
// subsystem particle numbers (N_A^up, N_A^dn), so the library forms rho_A per block (FP64
// MFMA) and returns S_vN = -sum lambda ln lambda together with the reference's
// S_ref = -sum sigma ln sigma (sigma = singular values).  Asite lists the sites of A.
#pragma once

#include "ed.hpp"

namespace ed {

template <typename T>
class EEmatHubAB_GPU {
  public:
    EEmatHubAB_GPU(const int *Asite, int Asize, const H_TBq &up, const H_TBq &dn)
        : N(up.N), Qup(up.Q), Qdn(dn.Q), dimUP(up.count), dimDN(dn.count)
    {
        for (int k = 0; k < Asize; ++k) maskA |= int64_t(1) << Asite[k];
    }
    struct Entropy {
        double S_vN, S_ref;
        int64_t n_schmidt;
    };
    Entropy entropy(const device_vector<T> &psi) const
    {
        if (psi.size() != (size_t)(dimUP * dimDN)) throw Error(ED_EINVAL, "EEmatHubAB_GPU: psi has the wrong size");
        double out[2];
        int64_t nsv = 0;
        check(ed_entanglement_entropy(N, Qup, Qdn, maskA, is_cplx<T>::value ? 1 : 0, psi.data(), out, &nsv, nullptr),
              "EEmatHubAB_GPU::entropy");
        return Entropy{out[0], out[1], nsv};
    }
    // all eigenvalues of rho_A, descending
    std::vector<double> spectrum(const device_vector<T> &psi) const
    {
        int64_t n = 0;
        check(ed_entanglement_spectrum(N, Qup, Qdn, maskA, is_cplx<T>::value ? 1 : 0, psi.data(), nullptr, 0, &n,
                                       nullptr),
              "EEmatHubAB_GPU::spectrum");
        std::vector<double> lam((size_t)n);
        check(ed_entanglement_spectrum(N, Qup, Qdn, maskA, is_cplx<T>::value ? 1 : 0, psi.data(), lam.data(), n, &n,
                                       nullptr),
              "EEmatHubAB_GPU::spectrum");
        return lam;
    }
    int N, Qup, Qdn;
    int64_t dimUP, dimDN, maskA = 0;
};

} // namespace ed

// ed/totalspin.hpp -- TotalSpin::Hubbard (MBL_test/include/HubMatrix_TotalSpin.cuh:126-173):
// y = 1/2 (S+S- + S-S+) x in the Hubbard sector basis.  S(S+1) = (Q_up - Q_dn)^2/4 + <x|y>
// for a normalised x (ConservedTotalSpin_Kagome_randomHop.cu:121-123).
#pragma once

#include "ed.hpp"

namespace ed {
namespace TotalSpin {

class Hubbard {
  public:
    Hubbard(const H_TBq &HqUP, const H_TBq &HqDN) : dimUP(HqUP.count), dimDN(HqDN.count), N(HqUP.N)
    {
        check(ed_s2_create(N, dimUP, HqUP.d_state.data(), HqUP.d_address.data(), dimDN, HqDN.d_state.data(),
                           HqDN.d_address.data(), &s),
              "TotalSpin::Hubbard");
    }
    Hubbard(const Hubbard &) = delete;
    Hubbard &operator=(const Hubbard &) = delete;
    ~Hubbard() { (void)ed_s2_destroy(s); }

    template <typename T>
    void Hubbard_mv(const device_vector<T> &x, device_vector<T> &y) const
    {
        if (x.size() != (size_t)(dimUP * dimDN) || y.size() != x.size())
            throw Error(ED_EINVAL, "TotalSpin::Hubbard_mv: vectors must hold dimUP*dimDN elements");
        check(ed_s2_apply(s, is_cplx<T>::value ? 1 : 0, x.data(), y.data(), nullptr), "TotalSpin::Hubbard_mv");
        hip_check(hipDeviceSynchronize(), "TotalSpin::Hubbard_mv");
    }

    int64_t dimUP, dimDN;
    int N;

  private:
    ed_s2 s = nullptr;
};

} // namespace TotalSpin
} // namespace ed

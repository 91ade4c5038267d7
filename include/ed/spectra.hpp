// ed/spectra.hpp -- cuSpectra::Hubbard<T> (MBL_test/include/HubMatrix_TotalSpin.cuh:181-322) and
// cuSpectra::SymEigsSolver<OpType> (SymEigsSolver.cuh:36-475, ArnoldiOp.cuh:132-219) with
// SortRule::SmallestAlge.  perform_op is the host-array H.v (device buffers allocated and
// copied per call, as in the reference); the solver itself runs on the device
// (ed_eigs_lowest: thick-restart Lanczos, full reorthogonalisation) and only needs the
// operator's handle.  Eigenvalues ascending; eigenvectors() returns nev columns of D elements
// (column k in elements [k*D, (k+1)*D)), on the host like Spectra's.
#pragma once

#include "ed.hpp"

namespace ed {
namespace cuSpectra {

enum class SortRule { SmallestAlge };
enum class CompInfo { Successful, NotComputed, NotConverging, NumericalIssue };

template <typename T>
class Hubbard {
  public:
    using Scalar = T;
    Hubbard(const H_TBq &HqUP, const H_TBq &HqDN, double U = 0.0) : H(HqUP, HqDN, U) {}
    int64_t rows() const { return H.dimUP * H.dimDN; }
    int64_t cols() const { return rows(); }
    void perform_op(const T *x_in, T *y_out) const
    {
        check(ed_hubbard_apply_host(H.handle(), is_cplx<T>::value ? 1 : 0, x_in, y_out, nullptr), "perform_op");
    }
    const ed::Hubbard &op() const { return H; }

  private:
    ed::Hubbard H;
};

template <typename OpType>
class SymEigsSolver {
  public:
    using T = typename OpType::Scalar;
    SymEigsSolver(OpType &op, int nev, int ncv) : op_(op), nev_(nev), ncv_(ncv)
    {
        if (nev < 1 || nev >= op.rows() || ncv <= nev) throw Error(ED_EINVAL, "SymEigsSolver: need 1 <= nev < ncv");
        if (ncv_ > op.rows()) ncv_ = (int)op.rows();
    }
    void init(unsigned seed = 235u) { seed_ = seed; }
    int compute(SortRule selection = SortRule::SmallestAlge, int maxit = 1000, double tol = 1e-10,
                SortRule sorting = SortRule::SmallestAlge)
    {
        (void)selection;
        (void)sorting;
        const int64_t D = op_.rows();
        device_vector<T> vecs((size_t)D * nev_);
        evals_.assign(nev_, 0.0);
        int nconv = 0, nit = 0;
        const int st = ed_eigs_lowest(op_.op().handle(), is_cplx<T>::value ? 1 : 0, nev_, ncv_, maxit, tol, seed_,
                                      evals_.data(), vecs.data(), &nconv, &nit, nullptr);
        if (st != ED_OK && st != ED_ENOCONV) check(st, "SymEigsSolver::compute");
        info_ = st == ED_OK ? CompInfo::Successful : CompInfo::NotConverging;
        hip_check(hipDeviceSynchronize(), "SymEigsSolver::compute");
        evecs_ = vecs.to_host();
        niter = nit;
        return nconv;
    }
    CompInfo info() const { return info_; }
    const std::vector<double> &eigenvalues() const { return evals_; }
    const std::vector<T> &eigenvectors() const { return evecs_; }
    int niter = 0;

  private:
    OpType &op_;
    int nev_, ncv_;
    unsigned seed_ = 235u;
    std::vector<double> evals_;
    std::vector<T> evecs_;
    CompInfo info_ = CompInfo::NotComputed;
};

} // namespace cuSpectra
} // namespace ed

// ed/hubmatrix.hpp -- Hubbard, Thrust_Lin<T>, Lanczos<T>, correlator_continued_fraction and
// compute_sigma_omega (MBL_test/include/HubMatrix_GPU.cuh:30-613, 816-849).
//
// Hubbard::Hubbard_mv(x, y) overwrites y = (H_up (x) 1 + 1 (x) H_dn + U popcount(up&dn&mask)) x
// with psi[i_up*dimDN + i_dn]; the Lanczos recurrence, its convergence test and the
// rebuild-by-recompute Ritz vector are those of Lanczos<T>::eigen, run on the device by
// ed_lanczos_eigen (alpha/beta are copied back into the host vectors a, b).
#pragma once

#include "ed.hpp"

namespace ed {

class Hubbard {
  public:
    // Hubbard(HqUP, HqDN, U) and Hubbard(HqUP, HqDN, U, dilution) (HubMatrix_GPU.cuh:46-95)
    Hubbard(const H_TBq &HqUP, const H_TBq &HqDN, double U_ = 0.0) : Hubbard(HqUP, HqDN, U_, -1) {}
    Hubbard(const H_TBq &HqUP, const H_TBq &HqDN, double U_, int dilution_)
        : dimUP(HqUP.count), dimDN(HqDN.count), N(HqUP.N), U(U_), is_diluted(dilution_ >= 0),
          dilution(dilution_ >= 0 ? dilution_ : (int)((int64_t(1) << HqUP.N) - 1))
    {
        if (HqUP.N != HqDN.N) throw Error(ED_EINVAL, "Hubbard: species with different N");
        if (HqUP.complex_tables() != HqDN.complex_tables())
            throw Error(ED_EINVAL, "Hubbard: both species need the same table type");
        cplx_tables = HqUP.complex_tables();
        const auto &cu = HqUP.calls(), &cd = HqDN.calls();
        if (HqUP.full_fock() || HqDN.full_fock()) {
            // full-Fock species: the table form (same operator)
            device_vector<double> axu, axd;
            device_vector<int32_t> aju, ajd;
            HqUP.tables(axu, aju);
            HqDN.tables(axd, ajd);
            check(ed_hubbard_create(N, dimUP, HqUP.slots(), axu.data(), aju.data(), HqUP.d_state.data(), dimDN,
                                    HqDN.slots(), axd.data(), ajd.data(), HqDN.d_state.data(), cplx_tables ? 1 : 0,
                                    U, is_diluted ? dilution : -1, nullptr, &h),
                  "Hubbard");
        } else {
            check(ed_hubbard_create_ops(N, HqUP.Q, cu.data(), (int)cu.size(), HqDN.Q, cd.data(), (int)cd.size(),
                                        cplx_tables ? 1 : 0, U, is_diluted ? dilution : -1, nullptr, &h),
                  "Hubbard");
        }
    }
    Hubbard(const Hubbard &) = delete;
    Hubbard &operator=(const Hubbard &) = delete;
    ~Hubbard() { (void)ed_hubbard_destroy(h); }

    // Hubbard::Hubbard_mv<M> (HubMatrix_GPU.cuh:163-182), M = double or std::complex<double>
    template <typename M>
    void Hubbard_mv(const device_vector<M> &x, device_vector<M> &y) const
    {
        if (x.size() != (size_t)(dimUP * dimDN) || y.size() != x.size())
            throw Error(ED_EINVAL, "Hubbard_mv: vectors must hold dimUP*dimDN elements");
        check(ed_hubbard_apply(h, is_cplx<M>::value ? 1 : 0, x.data(), y.data(), nullptr), "Hubbard_mv");
        hip_check(hipDeviceSynchronize(), "Hubbard_mv");
    }

    ed_hubbard handle() const { return h; }
    bool complex_tables() const { return cplx_tables; }

    // Hubbard::Write (HubMatrix_GPU.cuh:219-230): ADDS H into the (dimUP*dimDN)^2 row-major dense
    // matrix Hmat (complex tables: interleaved complex entries, 2x the doubles)
    void Write(device_vector<double> &Hmat) const
    {
        const size_t D = (size_t)(dimUP * dimDN);
        if (Hmat.size() != D * D * (cplx_tables ? 2 : 1)) throw Error(ED_EINVAL, "Hubbard::Write: Hmat size");
        check(ed_hubbard_write_dense(h, Hmat.data(), nullptr), "Hubbard::Write");
        hip_check(hipDeviceSynchronize(), "Hubbard::Write");
    }
    // Hubbard::Write_pinned (HubMatrix_GPU.cuh:232-248): Hmat is any device-accessible pointer
    void Write_pinned(double *Hmat) const
    {
        check(ed_hubbard_write_dense(h, Hmat, nullptr), "Hubbard::Write_pinned");
        hip_check(hipDeviceSynchronize(), "Hubbard::Write_pinned");
    }
    // Slab layout (no reference counterpart): builds the layout tables once; Lanczos<double>::eigen
    // then runs its recurrence on it (same eigenvalue and Ritz vector, 18-site H.v 72 -> 56 ms).
    // Returns the layout's vector length (0 after slab_layout(false)).
    int64_t slab_layout(bool enable = true)
    {
        int64_t n = 0;
        check(ed_hubbard_slab_layout(h, enable ? 1 : 0, &n), "Hubbard::slab_layout");
        return n;
    }
    // Hubbard::initialize (HubMatrix_GPU.cuh:257-277): zero the dense matrix
    void initialize(double *Hmat) const
    {
        check(ed_dense_zero(dimUP * dimDN, cplx_tables ? 1 : 0, Hmat, nullptr), "Hubbard::initialize");
        hip_check(hipDeviceSynchronize(), "Hubbard::initialize");
    }

    int64_t dimUP, dimDN;
    int N;
    double U;
    bool is_diluted;
    int dilution;

  private:
    ed_hubbard h = nullptr;
    bool cplx_tables = false;
};

// Thrust_Lin<T> (HubMatrix_GPU.cuh:291-471): BLAS-1 on device vectors.  norm() is the
// SQUARED norm, as in the reference.
template <typename T>
class Thrust_Lin {
  public:
    Thrust_Lin(int64_t dimUP_, int64_t dimDN_) : dimUP(dimUP_), dimDN(dimDN_) {}
    int64_t dimUP, dimDN;
    static constexpr int C = is_cplx<T>::value ? 1 : 0;

    T thrust_Dotprod(const device_vector<T> &v1, const device_vector<T> &v2) const
    {
        double out[2] = {0, 0};
        check(ed_dot((int64_t)v1.size(), C, v1.data(), v2.data(), out, nullptr), "thrust_Dotprod");
        if constexpr (is_cplx<T>::value) return T(out[0], out[1]);
        else return out[0];
    }
    double norm(const device_vector<T> &v) const { return std::real(thrust_Dotprod(v, v)); }
    void normalize(device_vector<T> &v) const
    {
        check(ed_normalize((int64_t)v.size(), C, v.data(), nullptr, nullptr), "normalize");
        hip_check(hipDeviceSynchronize(), "normalize");
    }
    // ax / axpy return new vectors, as the reference's do
    device_vector<T> ax(double s, const device_vector<T> &v) const
    {
        device_vector<T> w(v);
        check(ed_scal((int64_t)w.size(), C, s, w.data(), nullptr), "ax");
        return w;
    }
    device_vector<T> axpy(double s, const device_vector<T> &v1, const device_vector<T> &v2) const
    {
        device_vector<T> w(v2);
        check(ed_axpy((int64_t)w.size(), C, s, v1.data(), w.data(), nullptr), "axpy");
        return w;
    }
};

template <typename T>
class Lanczos : public Thrust_Lin<T> {
  public:
    using Thrust_Lin<T>::C;
    Lanczos(int64_t dimUP_, int64_t dimDN_) : Thrust_Lin<T>(dimUP_, dimDN_) { clear(); }

    std::vector<double> a, b;
    int maxN = 1000;
    double tolconv = 0.01 * std::sqrt(2.220446049250313e-16);
    int nsteps = 0;

    void clear() // HubMatrix_GPU.cuh:523-529
    {
        a.clear();
        b.assign(1, 0.0);
    }
    // GenRand(235) + normalize (HubMatrix_GPU.cuh:504-539)
    void random_vector(device_vector<T> &f) const
    {
        check(ed_random_vector((int64_t)f.size(), 235u, C, f.data(), 1, nullptr), "random_vector");
        hip_check(hipDeviceSynchronize(), "random_vector");
    }
    // Lanczos<T>::eigen (HubMatrix_GPU.cuh:561-613); f is clobbered, evec receives the Ritz
    // vector.  Throws ed::Error(ED_ENOCONV) when the test is not met in maxN steps.
    void eigen(const Hubbard &Hub, device_vector<T> &f, double &eval, device_vector<T> &evec)
    {
        if (evec.size() != f.size()) evec.resize(f.size());
        ed_lanczos_cfg cfg{maxN, tolconv, 10, 0};
        std::vector<double> ha(maxN + 2), hb(maxN + 3);
        int n = 0;
        const int st = ed_lanczos_eigen(Hub.handle(), C, &cfg, f.data(), &eval, evec.data(), ha.data(), hb.data(), &n,
                                        nullptr);
        a.assign(ha.begin(), ha.begin() + n);
        b.assign(hb.begin(), hb.begin() + n + 1);
        nsteps = n;
        check(st, "Lanczos::eigen");
    }
    // Lanczos<T>::tridiagonal (HubMatrix_GPU.cuh:541-559): appends nsteps a's and b's
    void tridiagonal(const Hubbard &Hub, const device_vector<T> &f, int nstep = 200)
    {
        std::vector<double> ha(nstep), hb(nstep + 1);
        check(ed_lanczos_tridiagonal(Hub.handle(), C, nstep, f.data(), ha.data(), hb.data(), nullptr),
              "Lanczos::tridiagonal");
        b[0] = hb[0];
        a.insert(a.end(), ha.begin(), ha.end());
        b.insert(b.end(), hb.begin() + 1, hb.end());
    }
};

// Krylov<C> (HubMatrix_GPU.cuh:619-728): real-time evolution exp(-i H dt) phi0 in the Lanczos
// space of phi0 (complex vectors).  a, b (b starts as {0}), m, z are host vectors as in the
// reference; the recurrence stops after the first b < btol or maxN steps.
template <typename Cr>
class Krylov : public Thrust_Lin<std::complex<Cr>> {
  public:
    using V = device_vector<std::complex<Cr>>;
    Krylov(int64_t dimUP_, int64_t dimDN_) : Thrust_Lin<std::complex<Cr>>(dimUP_, dimDN_) { clear(); }

    std::vector<double> a, b, m, z;
    int maxN = 1000;
    double btol = 1e-10;

    void clear() // HubMatrix_GPU.cuh:730-739
    {
        a.clear();
        b.assign(1, 0.0);
        m.clear();
        z.clear();
    }
    // Krylov<C>::tridiagonal (HubMatrix_GPU.cuh:660-680): f is used as given
    void tridiagonal(const Hubbard &Hub, const V &f)
    {
        std::vector<double> ha(maxN), hb(maxN + 1);
        int n = 0;
        check(ed_krylov_tridiagonal(Hub.handle(), f.data(), maxN, btol, ha.data(), hb.data(), &n, nullptr),
              "Krylov::tridiagonal");
        a.assign(ha.begin(), ha.begin() + n);
        b.assign(hb.begin(), hb.begin() + n + 1);
    }
    // HubMatrix_GPU.cuh:683-692
    void Make_tridiagonal_eigen(const Hubbard &Hub, const V &phi0)
    {
        tridiagonal(Hub, phi0);
        const int n = (int)a.size();
        m.resize(n);
        z.resize((size_t)n * n);
        std::vector<double> e(b.begin() + 1, b.begin() + std::max(1, n));
        if (e.empty()) e.push_back(0.0);
        check(ed_tridiag_eigen(n, a.data(), e.data(), m.data(), z.data()), "Krylov::Make_tridiagonal_eigen");
    }
    // Krylov<C>::Krylov_time_evol (HubMatrix_GPU.cuh:694-724): phi accumulates
    void Krylov_time_evol(const Hubbard &Hub, const V &phi0, V &phi, double dt)
    {
        check(ed_krylov_time_evol(Hub.handle(), phi0.data(), (int)a.size(), a.data(), b.data(), m.data(), z.data(),
                                  dt, phi.data(), nullptr),
              "Krylov::Krylov_time_evol");
        hip_check(hipDeviceSynchronize(), "Krylov::Krylov_time_evol");
    }
};

// cusolverDsyevd (EIGEN_cusolver.h:16-63): eigenvalues ascending into d_W, eigenvectors over
// d_A (vector k contiguous at d_A + k*m); rocSOLVER syevd
inline void cusolverDsyevd(const int m, double *d_A, double *d_W)
{
    check(ed_dense_eigh(m, 0, d_A, d_W, nullptr), "cusolverDsyevd");
}
// cusolverZheevd (EIGEN_cusolver.h:65-112), complex128 Hermitian
inline void cusolverZheevd(const int m, std::complex<double> *d_A, double *d_W)
{
    check(ed_dense_eigh(m, 1, d_A, d_W, nullptr), "cusolverZheevd");
}

// tED<C> (HubMatrix_GPU.cuh:730-814): propagation in the eigenbasis of the dense Hamiltonian.
// W holds the eigenvalues, V the eigenvectors (V[k*D ...] is eigenvector k), both on the device.
template <typename Cr>
class tED : public Thrust_Lin<Cr> {
  public:
    tED(int64_t dimUP_, int64_t dimDN_) : Thrust_Lin<Cr>(dimUP_, dimDN_), W((size_t)(dimUP_ * dimDN_)) {}
    device_vector<double> W;
    device_vector<Cr> V;

    // Make_eigen (HubMatrix_GPU.cuh:756-767): V = Hmat, then syevd
    void Make_eigen(const device_vector<Cr> &Hmat)
    {
        V = Hmat;
        const int64_t D = this->dimUP * this->dimDN;
        if (W.size() != (size_t)D) W.resize((size_t)D);
        check(ed_dense_eigh(D, 0, V.data(), W.data(), nullptr), "tED::Make_eigen");
    }
    // tED_time_evol (HubMatrix_GPU.cuh:777-790): phi += sum_k e^{-i W_k dt} <V_k|phi0> V_k
    void tED_time_evol(const device_vector<std::complex<Cr>> &phi0, device_vector<std::complex<Cr>> &phi, double dt)
    {
        const int64_t D = this->dimUP * this->dimDN;
        check(ed_ted_time_evol(D, 0, V.data(), W.data(), phi0.data(), dt, phi.data(), nullptr), "tED::tED_time_evol");
    }
    void clear() // HubMatrix_GPU.cuh:807-812
    {
        W.resize((size_t)(this->dimUP * this->dimDN));
        V.clear();
    }
};

// correlator_continued_fraction (HubMatrix_GPU.cuh:820-835): phi is normalised in place
template <typename T>
void correlator_continued_fraction(Lanczos<T> &lanctri, const Hubbard &H, int nw,
                                   const std::vector<std::complex<double>> &w, const double &energy,
                                   device_vector<T> &phi, std::vector<std::complex<double>> &Lambda)
{
    const double norm = lanctri.norm(phi);
    lanctri.normalize(phi);
    lanctri.tridiagonal(H, phi);
    Lambda.resize(nw);
    check(ed_continued_fraction((int)lanctri.a.size(), lanctri.a.data(), lanctri.b.data(), energy, norm, nw,
                                reinterpret_cast<const double *>(w.data()), reinterpret_cast<double *>(Lambda.data())),
          "correlator_continued_fraction");
}

// compute_sigma_omega (HubMatrix_GPU.cuh:837-849): Lambda(w) of the correlator of O on evec
template <typename T>
void compute_sigma_omega(Lanczos<T> &lanctri, const Hubbard &H, const Hubbard &O, const double &E,
                         const device_vector<T> &evec, const std::vector<std::complex<double>> &w,
                         std::vector<std::complex<double>> &Lambda)
{
    lanctri.clear();
    device_vector<T> phi(evec.size());
    O.Hubbard_mv(evec, phi);
    correlator_continued_fraction(lanctri, H, (int)w.size(), w, E, phi, Lambda);
}

} // namespace ed

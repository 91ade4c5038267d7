// ed/fermi_operator.hpp -- H_TBq (MBL_test/include/fermi_operator.h:42-242): one spin
// species' basis (all N-bit integers with popcount Q, ascending = colex rank) and its operator
// call list.  The reference fills ELL tables on the host inside every Make_* call; here the
// calls are recorded in order and the library either generates the hops on the fly
// (ed_hubbard_create_ops) or builds the tables on the device (tables(), ed_tables_build) --
// replaying the calls in order gives the reference's slot placement exactly.
#pragma once

#include <map>

#include "ed.hpp"

namespace ed {

// fermi_operator.h:19-24
inline int count1(int64_t state)
{
    int c = 0;
    for (; state; ++c) state &= state - 1;
    return c;
}

// fermi_operator.h:26-40: occupied sites strictly between i and j
inline int count1_bt_sites(int64_t basis, int i_site, int j_site)
{
    const int lo = i_site < j_site ? i_site : j_site, hi = i_site < j_site ? j_site : i_site;
    if (hi - lo < 2) return 0;
    return count1((basis >> (lo + 1)) & ((int64_t(1) << (hi - lo - 1)) - 1));
}

class H_TBq {
  public:
    // H_TBq(N, Q) (fermi_operator.h:80-88); complex_tables enables Make_H_phase
    H_TBq(const int &N_, const int &Q_, bool complex_tables = false) : N(N_), Q(Q_), cplx(complex_tables)
    {
        init(Q_);
    }
    // H_TBq(N): the full Fock space (fermi_operator.h:90-98)
    explicit H_TBq(const int &N_) : N(N_), Q(N_), full(true) { init(-1); }

    int N, Q;
    int64_t dim = 0, count = 0;
    std::vector<int> state, address; // host copies (drivers index Hq.address[...] directly)
    device_vector<int32_t> d_state, d_address;

    void Make_H(int i, int j, double t) { ops.push_back(ed_op{ED_OP_HOP, i, j, 0, t, 0.0}); }     // :100-115
    void Make_J(int i, int j, double t) { ops.push_back(ed_op{ED_OP_CURRENT, i, j, 0, t, 0.0}); } // :117-138
    void Make_N(int i, double v) { ops.push_back(ed_op{ED_OP_DENSITY, i, i, 0, v, 0.0}); }        // :140-149
    void Insert_2Nop(int i, int j, double v) { ops.push_back(ed_op{ED_OP_DENSITY2, i, j, 0, v, 0.0}); } // :216-225
    // t e^{i phase} c+_i c_j + h.c. (Peierls twist; complex tables only, no reference counterpart)
    void Make_H_phase(int i, int j, double t, double phase)
    {
        if (!cplx) throw Error(ED_EINVAL, "Make_H_phase needs complex tables");
        ops.push_back(ed_op{ED_OP_HOP_PHASE, i, j, 0, t, phase});
    }
    // Insert_1Fop(i_dag, i, t) (fermi_operator.h:151-174): t c^{i_dag}_i, full Fock space only
    void Insert_1Fop(const char i_dag, int i, double t)
    {
        ops.push_back(ed_op{ED_OP_FERMI1, i, i, dag_bit(i_dag), t, 0.0});
    }
    // Insert_2Fop(i_dag, i, j_dag, j, t) (fermi_operator.h:176-214)
    void Insert_2Fop(const char i_dag, int i, const char j_dag, int j, double t)
    {
        ops.push_back(ed_op{ED_OP_FERMI2, i, j, dag_bit(i_dag) | (dag_bit(j_dag) << 1), t, 0.0});
    }
    void clear() { ops.clear(); } // fermi_operator.h:237-241 (basis kept)

    // H_TBq::Write (fermi_operator.h:227-235): Mat[j*count + Aj] += Ax over the nonzero slots;
    // Mat holds count*count doubles (real tables)
    void Write(device_vector<double> &Mat) const
    {
        if (cplx) throw Error(ED_EINVAL, "H_TBq::Write: real tables only");
        if (Mat.size() != (size_t)(count * count)) throw Error(ED_EINVAL, "H_TBq::Write: Mat must hold count^2");
        device_vector<double> Ax;
        device_vector<int32_t> Aj;
        tables(Ax, Aj);
        check(ed_tables_write_dense(count, slots(), 0, Ax.data(), Aj.data(), Mat.data(), nullptr), "H_TBq::Write");
        hip_check(hipDeviceSynchronize(), "H_TBq::Write");
    }

    bool complex_tables() const { return cplx; }
    bool full_fock() const { return full; }
    int slots() const { return 10 * Q + 1; }
    const std::vector<ed_op> &calls() const { return ops; }

    // The reference's ELL tables (column-major, slot k of row a at k*count + a), built on the
    // device from the call list: Ax has count*slots doubles (x2 interleaved when complex).
    void tables(device_vector<double> &Ax, device_vector<int32_t> &Aj) const
    {
        const size_t n = (size_t)count * slots();
        if (Ax.size() != n * (cplx ? 2 : 1)) Ax.resize(n * (cplx ? 2 : 1));
        if (Aj.size() != n) Aj.resize(n);
        check(ed_tables_build(N, full ? -1 : Q, d_state.data(), d_address.data(), count, ops.data(), (int)ops.size(),
                              cplx ? 1 : 0, Ax.data(), Aj.data(), nullptr),
              "H_TBq::tables");
        hip_check(hipDeviceSynchronize(), "H_TBq::tables");
    }

  private:
    static int dag_bit(char c)
    {
        if (c == '+') return 0;
        if (c == '-') return 1;
        throw Error(ED_EINVAL, "H_TBq: operator sign must be '+' or '-'"); // the reference prints "wrong sign!"
    }
    void init(int q)
    {
        dim = int64_t(1) << N;
        check(ed_basis_count(N, q, &count), "H_TBq: basis count");
        d_state.resize((size_t)count);
        d_address.resize((size_t)dim);
        check(ed_basis_build(N, q, d_state.data(), d_address.data(), nullptr), "H_TBq: basis");
        hip_check(hipDeviceSynchronize(), "H_TBq: basis");
        std::vector<int32_t> s = d_state.to_host(), a = d_address.to_host();
        state.assign(s.begin(), s.end());
        address.assign(a.begin(), a.end());
    }
    bool cplx = false, full = false;
    std::vector<ed_op> ops;
};

namespace LOW_MEM {
// LOW_MEM::H_TBq basis (TotalSpin/include/fermi_operator_lowmem.h:50-64): 64-bit states in
// prev_permutation order; address is the rank map the reference keeps in a std::map.
struct H_TBq_basis {
    H_TBq_basis(int N_, int Q_) : N(N_), Q(Q_)
    {
        int64_t n = 0;
        if (Q_ < 0 || Q_ > N_ || N_ > 63) throw Error(ED_EINVAL, "LOW_MEM::H_TBq: bad N, Q");
        long double c = 1;
        for (int k = 0; k < Q_; ++k) c = c * (N_ - k) / (k + 1);
        n = (int64_t)(c + 0.5L);
        device_vector<uint64_t> d((size_t)n);
        check(ed_basis_build_lowmem(N_, Q_, d.data(), nullptr), "LOW_MEM::H_TBq");
        hip_check(hipDeviceSynchronize(), "LOW_MEM::H_TBq");
        state = d.to_host();
        count = (int)n;
        for (int a = 0; a < count; ++a) address[state[a]] = a;
    }
    int N, Q, count = 0;
    std::vector<uint64_t> state;
    std::map<uint64_t, int> address;
};
} // namespace LOW_MEM

} // namespace ed

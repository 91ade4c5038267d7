// ref_fermi_driver.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Thin extern "C" driver around the reference's own host-side basis and table
// builder, MBL_test/include/fermi_operator.h (H_TBq, fermi_operator.h:42-242),
// included from the reference tree at build time (oracle/Makefile, target `ref`).
// The built library lands in oracle/_ref/ and is used by the CPU tests to pin the
// oracle's basis/table restatement bit-exactly against the reference code.  No
// reference source is copied into this repository.
#include "fermi_operator.h"

#include <cstdint>
#include <cstring>

extern "C" {

// Builds an H_TBq(N,Q) (Q < 0: the full-Fock H_TBq(N)), applies the operator list in
// order and copies out state[count], address[2^N], Ax/Aj[count*(10*Q'+1)].
// kind: 0 Make_H, 1 Make_J, 2 Make_N, 3 Insert_2Nop, 5 Insert_1Fop, 6 Insert_2Fop; dag (may be
// NULL) bit 0 / bit 1 select '-' (1) or '+' (0) for the operator on site i / j.  Returns count.
long long ref_htbq_tables(int N, int Q, int nops, const int *kind, const int *si, const int *sj, const double *val,
                          int *state_out, int *address_out, double *Ax_out, int *Aj_out, const int *dag)
{
    H_TBq *H = (Q < 0) ? new H_TBq(N) : new H_TBq(N, Q);
    for (int o = 0; o < nops; ++o) {
        const int dg = dag ? dag[o] : 0;
        const char ci = (dg & 1) ? '-' : '+', cj = (dg & 2) ? '-' : '+';
        switch (kind[o]) {
        case 0: H->Make_H(si[o], sj[o], val[o]); break;
        case 1: H->Make_J(si[o], sj[o], val[o]); break;
        case 2: H->Make_N(si[o], val[o]); break;
        case 3: H->Insert_2Nop(si[o], sj[o], val[o]); break;
        case 5: H->Insert_1Fop(ci, si[o], val[o]); break;
        case 6: H->Insert_2Fop(ci, si[o], cj, sj[o], val[o]); break;
        default: break;
        }
    }
    long long count = H->count;
    for (long long a = 0; a < count; ++a) state_out[a] = H->state[a];
    for (long long n = 0; n < (1LL << N); ++n) address_out[n] = H->address[n];
    for (size_t k = 0; k < H->Ax.size(); ++k) {
        Ax_out[k] = H->Ax[k];
        Aj_out[k] = H->Aj[k];
    }
    delete H;
    return count;
}

int ref_count1_bt_sites(int basis, int i, int j) { return count1_bt_sites(basis, i, j); }
int ref_count1(int s) { return count1(s); }
}

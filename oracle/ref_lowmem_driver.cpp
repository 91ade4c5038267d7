// ref_lowmem_driver.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Thin extern "C" driver around the reference's 64-bit basis builder,
// TotalSpin/include/fermi_operator_lowmem.h (LOW_MEM::H_TBq, Labeling at :50-64), included from
// the reference tree at build time (oracle/Makefile, target `ref`).  The built library lands in
// oracle/_ref/ and pins the oracle's and the library's LOW_MEM basis order in the CPU tests.
#include "fermi_operator_lowmem.h"

#include <cstdint>

extern "C" {

// LOW_MEM::H_TBq(N, Q).state copied into out (capacity cap); returns count.
long long ref_lowmem_basis(int N, int Q, unsigned long long *out, long long cap)
{
    LOW_MEM::H_TBq H(N, Q);
    const long long n = H.count;
    for (long long a = 0; a < n && a < cap; ++a) out[a] = H.state[a];
    return n;
}
}

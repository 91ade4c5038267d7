"""MI355X-native exact diagonalization of spinful fermion lattice models.

Host-side mirror of the reference's C++ interfaces (H_TBq, Hubbard, Lanczos,
Krylov and dense (tED) time evolution, the dense eigensolver, TotalSpin::Hubbard, the
continued-fraction correlator and the entanglement entropy) over the C ABI of libedhip.so (include/edhip.h).  Nothing here computes
on the CPU: every entry point calls the HIP library and fails loudly when it is
missing.
"""
from .fermi_operator import H_TBq, count1, count1_bt_sites, lowmem_basis  # noqa: F401
from .hubmatrix import (Hubbard, Krylov, Lanczos, Thrust_Lin, compute_sigma_omega,  # noqa: F401
                        correlator_continued_fraction, cusolverDsyevd, cusolverZheevd, drude_weight_Jx,
                        lanczos_orthogonality, tED, tridiag_eigen)
from . import lattices  # noqa: F401

__all__ = ["H_TBq", "Hubbard", "Krylov", "Lanczos", "Thrust_Lin", "compute_sigma_omega",
           "correlator_continued_fraction", "cusolverDsyevd", "cusolverZheevd", "drude_weight_Jx",
           "lanczos_orthogonality", "tED",
           "tridiag_eigen", "lattices", "count1", "count1_bt_sites", "lowmem_basis"]

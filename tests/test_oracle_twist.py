"""The Lieb-12 twist convention of the batched Drude sweep, pinned on the CPU oracle.

The sweep puts a total Peierls phase Phi around each x ring, spread as Phi/4 on its four bonds
(lattices.lieb_twisted_ops).  Two facts fix that convention against the reference:
- gauge equivalence: the spread twist and a boundary twist (the whole Phi on the last bond of
  each ring, the bond the reference flips) have the same spectrum at every Phi, while the
  spectrum itself does move with Phi;
- Phi = pi reproduces the reference's real pi-flux Lieb cluster, whose x bonds 3 and 7 carry +1
  instead of -1 (LiebLattice12, TotalSpinExcited_Lieb_randomHop.cu:176-188).
Spectra are compared in full by dense diagonalisation of small sectors and through the Lanczos
ground-state energy (Lanczos<T>::eigen, HubMatrix_GPU.cuh:561-613) of the 3/3 sector.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from exact_diagonalization_for_fermionic_systems_with_gpu_amd import lattices as L

U = 4.0
SMALL = [(1, 1), (2, 1), (1, 3)]


def spectrum(ops, nu, nd, cplx=True):
    H = O.Hubbard(O.Species(12, nu, ops, cplx), O.Species(12, nd, ops, cplx), U)
    return np.linalg.eigvalsh(H.dense())


@pytest.mark.parametrize("nu,nd", SMALL)
def test_boundary_twist_is_gauge_equivalent(nu, nd):
    e_prev = None
    for k in (1, 3, 5, 8):
        phi = 2 * math.pi * k / 16
        es = spectrum(L.lieb_twisted_ops(phi), nu, nd)
        eb = spectrum(L.lieb_twisted_ops(phi, spread=False), nu, nd)
        assert np.max(np.abs(es - eb)) <= 1e-12 * max(1.0, np.max(np.abs(es)))
        if e_prev is not None and nu + nd > 1:
            assert np.max(np.abs(es - e_prev)) > 1e-6  # the twist does move the levels
        e_prev = es


@pytest.mark.parametrize("nu,nd", SMALL)
def test_pi_twist_is_the_reference_pi_flux(nu, nd):
    ref = spectrum(L.lieb_pi_flipped_ops(), nu, nd, cplx=False)
    for spread in (True, False):
        e = spectrum(L.lieb_twisted_ops(math.pi, spread=spread), nu, nd)
        assert np.max(np.abs(e - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))
    # and it is not the untwisted cluster
    e0 = spectrum(L.lieb_twisted_ops(0.0), nu, nd)
    if nu + nd > 1:
        assert np.max(np.abs(e0 - ref)) > 1e-6


def test_pi_twist_ground_state_energy_3_3():
    """E0 of the 3/3 sector (dim 48,400) by the oracle Lanczos: spread and boundary twist at
    Phi = pi (complex tables) against the real pi-flux cluster, to the 1e-10 bar."""
    def e0(ops, cplx):
        H = O.Hubbard(O.Species(12, 3, ops, cplx), O.Species(12, 3, ops, cplx), U)
        st, e, *_ = H.lanczos_eigen(O.random_vector(H.dim, complex_=cplx))
        assert st == 0
        return e
    ref = e0(L.lieb_pi_flipped_ops(), False)
    for spread in (True, False):
        e = e0(L.lieb_twisted_ops(math.pi, spread=spread), True)
        assert abs(e - ref) <= 1e-10 * abs(ref), (spread, e, ref)

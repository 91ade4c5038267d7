"""Pins of the entanglement oracle (entangle_entropy.h:16-205, Hubbard_EE_test.cu:77-83).

* the particle-number-block decomposition reproduces the singular values of the reference's
  dense first-appearance-labelled matrix (so the block route is the reference's entropy);
* product state -> S = 0; two-site singlet -> S_vN = ln 2, S_ref = (sqrt2/2) ln 2;
* sum lambda = 1, S_A = S_B for a pure state, rho_A positive semidefinite.
"""
import math

import numpy as np
import pytest

from oracle import ee_oracle as E


def _rand_state(n, seed, cplx=False):
    r = np.random.default_rng(seed)
    v = r.normal(size=n) + (1j * r.normal(size=n) if cplx else 0)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("N,Qu,Qd,A,cplx", [(6, 3, 3, [0, 1, 2], False), (6, 2, 3, [0, 2, 4], False),
                                            (8, 4, 3, [0, 1, 2, 3], True), (7, 3, 3, [1, 2, 5], False)])
def test_blocks_equal_reference_matrix(N, Qu, Qd, A, cplx):
    dim = math.comb(N, Qu) * math.comb(N, Qd)
    psi = _rand_state(dim, N * 100 + Qu, cplx)
    sref, s = E.reference_entropy(N, Qu, Qd, A, psi)
    lam, sig = E.block_spectra(N, Qu, Qd, sum(1 << i for i in A), psi)
    k = len(sig)
    assert np.allclose(np.sort(sig)[::-1], np.sort(s)[::-1][:k], atol=1e-13)
    assert np.all(np.sort(s)[::-1][k:] < 1e-13)
    assert abs(E.entropy_from_sigma(sig) - sref) < 1e-12
    assert abs(np.sum(lam) - 1) < 1e-12
    assert np.min(lam) > -1e-14
    # S_A = S_B
    lamB, _ = E.block_spectra(N, Qu, Qd, ((1 << N) - 1) ^ sum(1 << i for i in A), psi)
    assert abs(E.entropy_from_lambda(lam) - E.entropy_from_lambda(lamB)) < 1e-12


def test_product_state_and_singlet():
    # product state: one basis state
    dim = math.comb(6, 3) ** 2
    psi = np.zeros(dim)
    psi[17] = 1.0
    lam, sig = E.block_spectra(6, 3, 3, 0b000111, psi)
    assert E.entropy_from_lambda(lam) == 0.0 and E.entropy_from_sigma(sig) == 0.0
    # two sites, one up + one down, singlet (|up,dn> - |dn,up>)/sqrt2: A = site 0
    # basis per species: state 1 (site 0) -> index 0, state 2 (site 1) -> index 1
    psi = np.zeros(4)
    psi[0 * 2 + 1] = 1 / math.sqrt(2)   # up on 0, dn on 1
    psi[1 * 2 + 0] = -1 / math.sqrt(2)  # up on 1, dn on 0
    lam, sig = E.block_spectra(2, 1, 1, 0b01, psi)
    assert abs(E.entropy_from_lambda(lam) - math.log(2)) < 1e-15
    assert abs(E.entropy_from_sigma(sig) - math.sqrt(2) / 2 * math.log(2)) < 1e-15
    sref, _ = E.reference_entropy(2, 1, 1, [0], psi)
    assert abs(sref - math.sqrt(2) / 2 * math.log(2)) < 1e-15

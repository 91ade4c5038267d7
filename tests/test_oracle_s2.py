"""The oracle's S^2 (the restatement of TotalSpin kernel::Hubbard_mv_, HubMatrix_TotalSpin.cuh:
20-66, which gives S_xy = (S+S- + S-S+)/2) pinned by what SU(2) fixes, independently of its code:

- S^2 = S_xy + S_z^2 with S_z = (Q_up - Q_dn)/2 (the drivers' formula,
  ConservedTotalSpin_Kagome_randomHop.cu:121-123) has eigenvalues S(S+1) only;
- each S appears with the multiplicity of spin-S multiplets among Nel = Q_up + Q_dn electrons on N
  sites, n_S = d(S) - d(S+1) with d(s) = C(N, Nel/2 + s) C(N, Nel/2 - s), in every sector with
  |S_z| <= S;
- S^2 is symmetric and commutes with an SU(2)-symmetric Hubbard H (oracle dense H, uniform hopping
  on a ring with a diagonal bond, any U).

A wrong exchange sign, a dropped pair or a wrong diagonal coefficient breaks the spectrum.
"""
from math import comb

import numpy as np
import pytest

from oracle import oracle as O


def s2_dense(N, qu, qd):
    up, dn = O.Species(N, qu), O.Species(N, qd)
    D = up.count * dn.count
    M = np.empty((D, D))
    e = np.zeros(D)
    for c in range(D):
        e[c] = 1.0
        M[:, c] = O.s2_apply(up, dn, e)
        e[c] = 0.0
    return M


def multiplets(N, nel):
    def d(s2):  # s2 = 2 S_z, integer, same parity as nel
        a, b = (nel + s2) // 2, (nel - s2) // 2
        return comb(N, a) * comb(N, b) if 0 <= a <= N and 0 <= b <= N else 0
    return {s2: d(s2) - d(s2 + 2) for s2 in range(nel % 2, nel + 1, 2) if d(s2) - d(s2 + 2) > 0}


SECTORS = [(4, 2, 2), (4, 3, 1), (5, 2, 2), (5, 3, 2), (6, 3, 3), (6, 4, 2), (6, 5, 1), (7, 4, 3), (7, 3, 3),
           (8, 4, 4)]


@pytest.mark.parametrize("N,qu,qd", SECTORS)
def test_s2_spectrum_is_su2(N, qu, qd):
    Sxy = s2_dense(N, qu, qd)
    assert np.max(np.abs(Sxy - Sxy.T)) == 0.0
    sz = (qu - qd) / 2.0
    ev = np.linalg.eigvalsh(Sxy + sz * sz * np.eye(len(Sxy)))
    expect = []
    for s2, m in multiplets(N, qu + qd).items():
        if s2 >= abs(qu - qd):
            S = s2 / 2.0
            expect += [S * (S + 1.0)] * m
    expect = np.sort(np.array(expect))
    assert len(expect) == len(ev)
    assert np.max(np.abs(ev - expect)) <= 1e-10


@pytest.mark.parametrize("N,qu,qd", [(6, 3, 3), (6, 4, 2), (7, 4, 3)])
def test_s2_commutes_with_hubbard(N, qu, qd):
    bonds = [(i, (i + 1) % N) for i in range(N)] + [(0, N // 2)]
    ops = [(0, i, j, -1.0) for i, j in bonds]
    H = O.Hubbard(O.Species(N, qu, ops), O.Species(N, qd, ops), 3.0).dense()
    S = s2_dense(N, qu, qd)
    assert np.max(np.abs(S @ H - H @ S)) <= 1e-12 * np.max(np.abs(H))
    # a hopping that breaks SU(2) (different for the two species) does not commute
    opd = [(0, i, j, -1.0 if k else -0.5) for k, (i, j) in enumerate(bonds)]
    H2 = O.Hubbard(O.Species(N, qu, ops), O.Species(N, qd, opd), 3.0).dense()
    assert np.max(np.abs(S @ H2 - H2 @ S)) > 1e-3

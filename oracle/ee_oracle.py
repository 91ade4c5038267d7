"""Entanglement-entropy oracle (numpy).  TEST INFRASTRUCTURE ONLY.

Two restatements, each citing the reference:

* reference_matrix(): the dense reshaped matrix of EEmatHubAB::Make_Write
  (MBL_test/include/entangle_entropy.h:16-205): per species, A- and B-substates are numbered
  in order of first appearance while scanning the sector basis (LabelingAB, :22-44); psi[i,j]
  goes to row (A_up*dimA_dn + A_dn), column (B_up*dimB_dn + B_dn) (:177-197).  Its singular
  values (numpy's LAPACK gesdd) give S_ref = -sum_{s != 0} s ln s as the drivers compute it
  (Hubbard_EE_test.cu:78-83).
* block_spectra(): the particle-number-block decomposition of the same matrix with colex
  numbering, rho_A = Psi Psi^H per block, eigenvalues by numpy eigvalsh and Schmidt values by
  numpy SVD.

Only small systems (dense SVD of the whole matrix) are meant for reference_matrix().
"""
from __future__ import annotations

import math

import numpy as np


def _basis(N, Q):
    return [n for n in range(1 << N) if bin(n).count("1") == Q]


def _labeling(states, A, B):
    """LabelingAB (entangle_entropy.h:22-44): first-appearance numbering of A and B parts."""
    addrA, addrB = {}, {}
    for s in states:
        a, b = s & A, s & B
        if a not in addrA:
            addrA[a] = len(addrA)
        if b not in addrB:
            addrB[b] = len(addrB)
    return addrA, addrB


def reference_matrix(N, Qu, Qd, Asites, psi):
    A = sum(1 << i for i in Asites)
    B = ((1 << N) - 1) ^ A
    su, sd = _basis(N, Qu), _basis(N, Qd)
    aU, bU = _labeling(su, A, B)
    aD, bD = _labeling(sd, A, B)
    rowD, colD = len(aU) * len(aD), len(bU) * len(bD)
    M = np.zeros((rowD, colD), dtype=np.asarray(psi).dtype)
    dd = len(sd)
    for i, u in enumerate(su):
        for j, d in enumerate(sd):
            row = aU[u & A] * len(aD) + aD[d & A]
            col = bU[u & B] * len(bD) + bD[d & B]
            M[row, col] = psi[i * dd + j]
    return M


def entropy_from_sigma(s):
    s = np.asarray(s)
    s = s[s != 0]
    return float(-np.sum(s * np.log(s)))


def entropy_from_lambda(lam):
    lam = np.asarray(lam)
    lam = lam[lam > 0]
    return float(-np.sum(lam * np.log(lam)))


def reference_entropy(N, Qu, Qd, Asites, psi):
    M = reference_matrix(N, Qu, Qd, Asites, psi)
    s = np.linalg.svd(M, compute_uv=False)
    return entropy_from_sigma(s), s


def _colex_rank(s):
    r, k = 0, 0
    p = 0
    while s:
        if s & 1:
            k += 1
            r += math.comb(p, k)
        s >>= 1
        p += 1
    return r


def _pext(s, mask):
    r, k = 0, 0
    for i in range(mask.bit_length()):
        if (mask >> i) & 1:
            if (s >> i) & 1:
                r |= 1 << k
            k += 1
    return r


def block_spectra(N, Qu, Qd, maskA, psi):
    """lambda (all rho_A eigenvalues) and sigma (all Schmidt values) from the blocks."""
    full = (1 << N) - 1
    maskB = full ^ maskA
    nA = bin(maskA).count("1")
    nB = N - nA
    psi = np.asarray(psi)
    su, sd = _basis(N, Qu), _basis(N, Qd)

    def codes(states):
        return [(bin(s & maskA).count("1"), _colex_rank(_pext(s, maskA)), _colex_rank(_pext(s, maskB)))
                for s in states]

    cu, cd = codes(su), codes(sd)
    blocks = {}
    dd = len(sd)
    for i, (au, rau, rbu) in enumerate(cu):
        for j, (ad, rad, rbd) in enumerate(cd):
            key = (au, ad)
            if key not in blocks:
                m = math.comb(nA, au) * math.comb(nA, ad)
                n = math.comb(nB, Qu - au) * math.comb(nB, Qd - ad)
                blocks[key] = np.zeros((m, n), dtype=psi.dtype)
            blocks[key][rau * math.comb(nA, ad) + rad, rbu * math.comb(nB, Qd - ad) + rbd] = psi[i * dd + j]
    lam, sig = [], []
    for M in blocks.values():
        rho = M @ M.conj().T if M.shape[0] <= M.shape[1] else M.conj().T @ M
        lam.extend(np.linalg.eigvalsh(rho).tolist())
        sig.extend(np.linalg.svd(M, compute_uv=False).tolist())
    return np.array(lam), np.array(sig)

"""Entanglement entropy of a Hubbard state for a bipartition A|B of the sites.

Mirror of EEmatHubAB_GPU (MBL_test/include/entangle_entropy_GPU.cuh:27-144) plus the
singular-value step the drivers run on it (cusolverDsvd/Zsvd, entangle_entropy.h:231-353;
S = -sum s ln s, Hubbard_EE_test.cu:77-83).  The reference materialises the whole
rowD x colD reshaped matrix; libedhip works block by block in (N_A^up, N_A^dn), which has
the same singular values (see ed_entangle.hip).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import call, dptr, stream_ptr
from .fermi_operator import H_TBq


class EEmatHubAB_GPU:
    """EEmatHubAB_GPU(Asite, Asize, HqUP, HqDN) (entangle_entropy_GPU.cuh:34-51)."""

    def __init__(self, Asite, Asize, HqUP: H_TBq, HqDN: H_TBq):
        self.Asite = [int(a) for a in list(Asite)[:Asize]]
        self.N, self.Qup, self.Qdn = HqUP.N, HqUP.Q, HqDN.Q
        self.dimUP, self.dimDN = HqUP.count, HqDN.count
        self.maskA = sum(1 << a for a in self.Asite)

    def entropy(self, psi: torch.Tensor):
        """(S_vN, S_ref, n_schmidt): -sum lambda ln lambda of rho_A, the reference's
        -sum sigma ln sigma, and the number of nonzero Schmidt values."""
        if psi.numel() != self.dimUP * self.dimDN:
            raise ValueError("psi has the wrong length")
        out = (C.c_double * 2)()
        nsv = C.c_int64()
        call("ed_entanglement_entropy", self.N, self.Qup, self.Qdn, self.maskA, 1 if psi.is_complex() else 0,
             dptr(psi), out, C.byref(nsv), stream_ptr())
        return out[0], out[1], nsv.value

    def spectrum(self, psi: torch.Tensor) -> np.ndarray:
        """all eigenvalues of rho_A, descending (one solve: rho_A has at most 4^min(|A|, |B|)
        eigenvalues, the number of (up, dn) configurations of the smaller side)"""
        if psi.numel() != self.dimUP * self.dimDN:
            raise ValueError("psi has the wrong length")
        nA = len(self.Asite)
        cap = 4 ** min(nA, self.N - nA)
        lam = np.zeros(cap)
        n = C.c_int64()
        call("ed_entanglement_spectrum", self.N, self.Qup, self.Qdn, self.maskA, 1 if psi.is_complex() else 0,
             dptr(psi), lam.ctypes.data_as(C.POINTER(C.c_double)), cap, C.byref(n), stream_ptr())
        if n.value > cap:
            raise RuntimeError("entanglement spectrum longer than its bound")
        return lam[: n.value]


def entropy_vn_batch(N: int, Q_up: int, Q_dn: int, maskA: int, states) -> np.ndarray:
    """S_vN of every state in `states` (tensors of one sector, all real or all complex) for the cut
    maskA, each block size of rho_A solved for all states at once
    (ed_entanglement_entropy_vn_batch)."""
    states = list(states)
    if not states:
        return np.zeros(0)
    cplx = states[0].is_complex()
    for s in states:
        if s.is_complex() != cplx or not s.is_contiguous() or not s.is_cuda:
            raise ValueError("entropy_vn_batch: states must be contiguous CUDA tensors of one dtype")
        if s.dtype not in (torch.float64, torch.complex128) or s.numel() != states[0].numel():
            raise ValueError("entropy_vn_batch: float64/complex128 states of one length")
    ptrs = (C.c_void_p * len(states))(*[s.data_ptr() for s in states])
    out = np.zeros(len(states))
    call("ed_entanglement_entropy_vn_batch", N, Q_up, Q_dn, maskA, 1 if cplx else 0, len(states), ptrs,
         out.ctypes.data_as(C.POINTER(C.c_double)), stream_ptr())
    return out

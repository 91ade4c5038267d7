"""cuSpectra interface: the Hubbard operator behind a host-pointer perform_op and the solver
for several lowest eigenpairs (MBL_test/include/HubMatrix_TotalSpin.cuh:181-322,
SymEigsSolver.cuh:36-475, ArnoldiOp.cuh:132-219).

    op = HubbardOp(H)                      # cuSpectra::Hubbard<double>(H_up, H_dn, U)
    eigs = SymEigsSolver(op, nev=4, ncv=20)
    eigs.init()
    nconv = eigs.compute("SmallestAlge", maxit=1000, tol=1e-10)
    eigs.info() == "Successful"; eigs.eigenvalues()

The solver runs on the device (ed_eigs_lowest: thick-restart Lanczos with full
reorthogonalisation, basis in HBM); perform_op is the reference's host-array H.v with device
buffers allocated and copied per call.  Eigenvalues come back in ascending order (the
reference's SmallestAlge sorting).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import ED_ENOCONV, call, lib, check, stream_ptr
from .hubmatrix import Hubbard


class HubbardOp:
    """cuSpectra::Hubbard<T>: rows(), cols(), perform_op(x_host, y_host)."""

    def __init__(self, H: Hubbard, complex_vectors: bool = False):
        self.H = H
        self.complex_vectors = bool(complex_vectors or H.complex_tables)
        self.Scalar = np.complex128 if self.complex_vectors else np.float64

    def rows(self) -> int:
        return self.H.dim

    def cols(self) -> int:
        return self.H.dim

    def perform_op(self, x_in: np.ndarray, y_out: np.ndarray):
        if x_in.dtype != self.Scalar or y_out.dtype != self.Scalar or x_in.size != self.rows() \
                or y_out.size != self.rows() or not x_in.flags.c_contiguous or not y_out.flags.c_contiguous:
            raise ValueError(f"perform_op: contiguous {self.Scalar.__name__} arrays of {self.rows()} elements needed")
        call("ed_hubbard_apply_host", self.H.handle, int(self.complex_vectors), x_in.ctypes.data_as(C.c_void_p),
             y_out.ctypes.data_as(C.c_void_p), stream_ptr())


class SymEigsSolver:
    """SymEigsSolver<OpType>(op, nev, ncv); init(); compute(...); info(); eigenvalues();
    eigenvectors()."""

    def __init__(self, op: HubbardOp, nev: int, ncv: int):
        n = op.rows()
        if not (1 <= nev < n) or not (nev < ncv):
            raise ValueError("need 1 <= nev < rows and nev < ncv (SymEigsSolver.cuh:448-475)")
        self.op, self.nev, self.ncv = op, int(nev), int(min(ncv, n))
        self.seed = 235
        self._evals = np.zeros(0)
        self._evecs = None
        self._info = "NotComputed"
        self.niter = 0

    def init(self, seed: int = 235):
        self.seed = int(seed)

    def compute(self, selection: str = "SmallestAlge", maxit: int = 1000, tol: float = 1e-10,
                sorting: str = "SmallestAlge", want_vectors: bool = True) -> int:
        if selection != "SmallestAlge" or sorting != "SmallestAlge":
            raise ValueError("only SortRule::SmallestAlge is used by the reference drivers")
        H = self.op.H
        ev = np.zeros(self.nev)
        nconv = C.c_int(0)
        nit = C.c_int(0)
        dt = torch.complex128 if self.op.complex_vectors else torch.float64
        vecs = torch.empty((self.nev, H.dim), dtype=dt, device="cuda") if want_vectors else None
        st = lib().ed_eigs_lowest(H.handle, int(self.op.complex_vectors), self.nev, self.ncv, int(maxit), float(tol),
                                  self.seed, ev.ctypes.data_as(C.POINTER(C.c_double)),
                                  C.c_void_p(vecs.data_ptr()) if vecs is not None else None, C.byref(nconv),
                                  C.byref(nit), stream_ptr())
        if st not in (0, ED_ENOCONV):
            check(st)
        self._info = "Successful" if st == 0 else "NotConverging"
        self._evals = ev
        self._evecs = vecs
        self.niter = nit.value
        return nconv.value

    def info(self) -> str:
        return self._info

    def eigenvalues(self) -> np.ndarray:
        return self._evals.copy()

    def eigenvectors(self) -> torch.Tensor:
        """[nev, D] device tensor (row k = eigenvector of eigenvalues()[k])."""
        return self._evecs

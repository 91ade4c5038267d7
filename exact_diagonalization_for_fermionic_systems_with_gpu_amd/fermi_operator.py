"""H_TBq -- single-species basis and ELL operator tables, on the GPU.

Mirror of class H_TBq (MBL_test/include/fermi_operator.h:42-242) with the same
method names and argument meaning.  The reference fills host tables inside each
Make_* call; here the calls are recorded in order and the tables are built by
one GPU kernel (ed_tables_build) the first time they are needed -- replaying the
calls in order gives the reference's slot placement exactly.

Public fields as in the reference: N, Q, dim (=2^N), count, state, address, Ax, Aj
(device tensors; .cpu() gives the host view the drivers index, e.g.
Hq.address[...] in Hubbard_EE_test.cu:49).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from ._lib import call, dptr, stream_ptr

Z_SLOTS = 10  # fermi_operator.h:11


def count1(state: int) -> int:
    """fermi_operator.h:19-24"""
    return bin(int(state)).count("1")


def count1_bt_sites(basis: int, i_site: int, j_site: int) -> int:
    """fermi_operator.h:26-40: occupied sites strictly between i and j"""
    lo, hi = min(i_site, j_site), max(i_site, j_site)
    return count1((int(basis) >> (lo + 1)) & ((1 << (hi - lo - 1)) - 1))


class H_TBq:
    """H_TBq(N, Q) (fermi_operator.h:80-88); H_TBq(N) is the full Fock space (90-98).

    complex_tables=True stores complex matrix elements (needed for Make_H_phase,
    the Peierls-twisted hop that has no counterpart in the reference)."""

    def __init__(self, N: int, Q: int | None = None, complex_tables: bool = False, device="cuda"):
        self.N = int(N)
        self.full = Q is None
        self.Q = self.N if Q is None else int(Q)
        self.dim = 1 << self.N
        self.complex_tables = bool(complex_tables)
        self.device = torch.device(device)
        cnt = C.c_int64()
        call("ed_basis_count", self.N, -1 if self.full else self.Q, C.byref(cnt))
        self.count = cnt.value
        self.slots = Z_SLOTS * self.Q + 1
        self.state = torch.empty(self.count, dtype=torch.int32, device=self.device)
        self.address = torch.empty(self.dim, dtype=torch.int32, device=self.device)
        call("ed_basis_build", self.N, -1 if self.full else self.Q, dptr(self.state), dptr(self.address),
             stream_ptr())
        self.ops: list[tuple] = []
        self._Ax = None
        self._Aj = None
        self._built = -1

    # --- the operator-insert methods (call order is kept) ------------------------------
    def Make_H(self, i_site, j_site, t):
        """fermi_operator.h:100-115"""
        self.ops.append((_lib.ED_OP_HOP, int(i_site), int(j_site), float(t), 0.0, 0))

    def Make_J(self, i_site, j_site, t):
        """fermi_operator.h:117-138"""
        self.ops.append((_lib.ED_OP_CURRENT, int(i_site), int(j_site), float(t), 0.0, 0))

    def Make_N(self, i_site, v):
        """fermi_operator.h:140-149"""
        self.ops.append((_lib.ED_OP_DENSITY, int(i_site), int(i_site), float(v), 0.0, 0))

    def Insert_2Nop(self, i_site, j_site, v):
        """fermi_operator.h:216-225"""
        self.ops.append((_lib.ED_OP_DENSITY2, int(i_site), int(j_site), float(v), 0.0, 0))

    def Make_H_phase(self, i_site, j_site, t, phase):
        """t e^{i phase} c+_i c_j + h.c. (complex tables only)."""
        if not self.complex_tables:
            raise ValueError("Make_H_phase needs H_TBq(..., complex_tables=True)")
        self.ops.append((_lib.ED_OP_HOP_PHASE, int(i_site), int(j_site), float(t), float(phase), 0))

    def Insert_1Fop(self, i_dag, i_site, t):
        """fermi_operator.h:151-174: t c^{i_dag}_i ('+' creates, '-' annihilates); full Fock
        space only, as the reference says"""
        self.ops.append((_lib.ED_OP_FERMI1, int(i_site), int(i_site), float(t), 0.0, _dag_bit(i_dag)))

    def Insert_2Fop(self, i_dag, i_site, j_dag, j_site, t):
        """fermi_operator.h:176-214: rows whose sites i, j are in the states the two operators
        act on; sign (-1)^(occupied sites strictly between i and j)"""
        if int(i_site) == int(j_site):
            raise ValueError("Insert_2Fop needs two different sites")
        self.ops.append((_lib.ED_OP_FERMI2, int(i_site), int(j_site), float(t), 0.0,
                         _dag_bit(i_dag) | (_dag_bit(j_dag) << 1)))

    def add_ops(self, ops):
        """ops: (kind, i, j, value[, phase[, dag]]) tuples"""
        for o in ops:
            self.ops.append((int(o[0]), int(o[1]), int(o[2]), float(o[3]), float(o[4]) if len(o) > 4 else 0.0,
                             int(o[5]) if len(o) > 5 else 0))

    def clear(self):
        """fermi_operator.h:237-241 (tables reset, basis kept)"""
        self.ops = []
        self._built = -1

    def ops_array(self):
        """the recorded calls as a C array of ed_op (for ed_tables_build / ed_hubbard_create_ops)"""
        arr = (_lib.EdOp * max(1, len(self.ops)))()
        for k, (kind, i, j, v, ph, dg) in enumerate(self.ops):
            arr[k] = _lib.EdOp(kind, i, j, dg, v, ph)
        return arr, len(self.ops)

    # --- tables ---------------------------------------------------------------------------
    def _build(self):
        if self._built == len(self.ops) and self._Ax is not None:
            return
        n = self.count * self.slots
        if self._Ax is None:
            self._Ax = torch.empty(n * (2 if self.complex_tables else 1), dtype=torch.float64, device=self.device)
            self._Aj = torch.empty(n, dtype=torch.int32, device=self.device)
        arr, _ = self.ops_array()
        call("ed_tables_build", self.N, -1 if self.full else self.Q, dptr(self.state), dptr(self.address),
             self.count, arr, len(self.ops), int(self.complex_tables), dptr(self._Ax), dptr(self._Aj),
             stream_ptr())
        self._built = len(self.ops)

    @property
    def Ax(self):
        self._build()
        if self.complex_tables:
            return torch.view_as_complex(self._Ax.view(-1, 2))
        return self._Ax

    @property
    def Aj(self):
        self._build()
        return self._Aj

    def raw_tables(self):
        self._build()
        return self._Ax, self._Aj

    def Write(self, Mat: torch.Tensor | None = None):
        """H_TBq::Write (fermi_operator.h:227-235): Mat[j, Aj] += Ax over the nonzero slots
        (ed_tables_write_dense).  Mat: a count x count float64 (complex128 for complex tables)
        device tensor, added into; None allocates a zeroed one.  Returns Mat."""
        Ax, Aj = self.raw_tables()
        dt = torch.complex128 if self.complex_tables else torch.float64
        if Mat is None:
            Mat = torch.zeros((self.count, self.count), dtype=dt, device=self.device)
        if Mat.numel() != self.count * self.count or Mat.dtype != dt or not Mat.is_contiguous():
            raise ValueError(f"Mat must be a contiguous {dt} tensor of count^2 elements")
        call("ed_tables_write_dense", self.count, self.slots, int(self.complex_tables), dptr(Ax), dptr(Aj),
             dptr(Mat), stream_ptr())
        return Mat


def _dag_bit(c) -> int:
    if c == "+":
        return 0
    if c == "-":
        return 1
    raise ValueError("operator sign must be '+' or '-'")  # the reference prints "wrong sign!"


def lowmem_basis(N: int, Q: int, device="cuda") -> torch.Tensor:
    """LOW_MEM::H_TBq basis (TotalSpin/include/fermi_operator_lowmem.h:50-64): the C(N,Q) states
    of N <= 63 sites as uint64 values (in an int64 tensor), in std::prev_permutation order
    (descending bit-reversed value).  The rank of a state is its index."""
    import math
    n = math.comb(int(N), int(Q))
    out = torch.empty(n, dtype=torch.int64, device=device)
    call("ed_basis_build_lowmem", int(N), int(Q), dptr(out), stream_ptr())
    return out

"""TotalSpin::Hubbard -- the (S+S- + S-S+)/2 operator (TotalSpin/include/HubMatrix_TotalSpin.cuh:126-173).

Drivers take Sxy = <V|y> and S(S+1) = (Q_up - Q_dn)^2/4 + Sxy
(ConservedTotalSpin_Kagome_randomHop.cu:121-123); spin_squared() does that.
"""
from __future__ import annotations

import ctypes as C

import torch

from ._lib import call, dptr, stream_ptr
from .fermi_operator import H_TBq


class Hubbard:
    def __init__(self, HqUP: H_TBq, HqDN: H_TBq):
        if HqUP.N != HqDN.N:
            raise ValueError("species on different lattices")
        self.N, self.dimUP, self.dimDN = HqUP.N, HqUP.count, HqDN.count
        self.Qup, self.Qdn = HqUP.Q, HqDN.Q
        h = C.c_void_p()
        call("ed_s2_create", self.N, self.dimUP, dptr(HqUP.state), dptr(HqUP.address), self.dimDN,
             dptr(HqDN.state), dptr(HqDN.address), C.byref(h))
        self._h = h

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                call("ed_s2_destroy", self._h)
                self._h = None
        except Exception:
            pass

    def Hubbard_mv(self, x: torch.Tensor, y: torch.Tensor):
        """HubMatrix_TotalSpin.cuh:157-173"""
        if x.numel() != self.dimUP * self.dimDN or y.numel() != x.numel() or x.dtype != y.dtype:
            raise ValueError("bad vector sizes")
        call("ed_s2_apply", self._h, 1 if x.is_complex() else 0, dptr(x), dptr(y), stream_ptr())

    def Hubbard_mv_new(self, x: torch.Tensor) -> torch.Tensor:
        y = torch.empty_like(x)
        self.Hubbard_mv(x, y)
        return y

    def spin_squared(self, V: torch.Tensor):
        """(Sxy, S(S+1)) of a normalized state V."""
        from .hubmatrix import Thrust_Lin
        y = self.Hubbard_mv_new(V)
        sxy = Thrust_Lin.thrust_Dotprod(V, y)
        sxy = sxy.real if isinstance(sxy, complex) else sxy
        return sxy, 0.25 * (self.Qup - self.Qdn) ** 2 + sxy


TotalSpinHubbard = Hubbard

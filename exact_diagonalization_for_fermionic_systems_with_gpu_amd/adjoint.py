"""Differentiable Hubbard H(theta) x and Hellmann-Feynman ground-state gradients.

Parameters theta = (t_b per bond, U_i per site, Phi twist flux) of

    H = sum_b sum_sigma -t_b (e^{i phi_b} c+_{j_b sigma} c_{i_b sigma} + h.c.) + sum_i U_i n_{i up} n_{i dn},
    phi_b = Phi * d_x(b) / L_x,

where bond b runs i_b -> j_b (the reference's first -> second, whose x projection d_x(b) the
Kubo current uses, drude_v2_cal_Kagome_randomHop.cu:841-870) and L_x is the x extent of the
cluster (one x winding picks up Phi).  The hop part is built as H_TBq calls (Make_H with a
Peierls phase), so H is applied by the same kernels as the reference-interface Hubbard.

HubbardApply is a torch.autograd.Function: forward y = H(theta) x; backward
grad_x = H grad_y (H is Hermitian) and grad_theta = Re <grad_y| dH/dtheta |x> from the
bilinear kernels ed_bond_bilinear / ed_site_double_occupancy.  ground_state_gradients runs
Lanczos (the Ritz vector rebuilt by re-running the recurrence, not stored) and returns
dE0/dtheta = <psi| dH/dtheta |psi>.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from ._lib import call, dptr, stream_ptr
from .fermi_operator import H_TBq
from .hubmatrix import Hubbard, Lanczos
from . import lattices as L


class ParamHubbard:
    """Builds the Hubbard operator for parameters (t[nb], U[N], phi) on a cluster."""

    def __init__(self, cluster: L.Cluster, Qup: int, Qdn: int, Lx: float):
        self.cl, self.Qup, self.Qdn, self.Lx = cluster, Qup, Qdn, float(Lx)
        self.dx = np.array([pj for (_, _, _, pj) in cluster.x_bonds]) if cluster.x_bonds else \
            np.zeros(len(cluster.bonds))
        self.nb = len(cluster.bonds)

    def ops(self, t, phi):
        # hop i -> j (term c+_j c_i) carries e^{+i phi_b}: call (j, i) with phase phi_b
        return [(L.OP_HOP_PHASE, j, i, -float(t[k]), float(phi * self.dx[k] / self.Lx))
                for k, (i, j) in enumerate(self.cl.bonds)]

    def build(self, t, U, phi):
        ops = self.ops(t, phi)
        N = self.cl.n_sites
        hu = H_TBq(N, self.Qup, complex_tables=True)
        hu.add_ops(ops)
        hd = H_TBq(N, self.Qdn, complex_tables=True)
        hd.add_ops(ops)
        return Hubbard(hu, hd, 0.0, Uvec=np.asarray(U, dtype=np.float64))

    def grads(self, H: Hubbard, a: torch.Tensor, b: torch.Tensor, t, phi):
        """Re <a| dH/dtheta |b> for theta = (t[nb], U[N], phi)."""
        cv = 1 if a.is_complex() else 0
        F = np.zeros((2, self.nb, 4))
        for sp in (0, 1):
            out = np.zeros(self.nb * 4)
            call("ed_bond_bilinear", H.handle, sp, cv, dptr(a), dptr(b), out.ctypes.data_as(C.POINTER(C.c_double)),
                 stream_ptr())
            F[sp] = out.reshape(self.nb, 4)
        Fc = (F[..., 0] + 1j * F[..., 1]).sum(0)  # sum over species
        Gc = (F[..., 2] + 1j * F[..., 3]).sum(0)
        value = -np.asarray(t, dtype=np.float64)
        g_t = -np.real((Fc + Gc) / value)                      # d/dt_b, value = -t_b
        g_phi = float(np.sum(np.real(1j * (Fc - Gc)) * self.dx / self.Lx))
        docc = np.zeros(self.cl.n_sites)
        call("ed_site_double_occupancy", H.handle, cv, dptr(a), dptr(b), docc.ctypes.data_as(C.POINTER(C.c_double)),
             stream_ptr())
        return g_t, docc, g_phi


class HubbardApply(torch.autograd.Function):
    """y = H(t, U, phi) x with gradients for x, t, U and phi."""

    @staticmethod
    def forward(ctx, x, t, U, phi, model: ParamHubbard):
        tn, Un, pn = t.detach().cpu().numpy(), U.detach().cpu().numpy(), float(phi.detach().cpu())
        H = model.build(tn, Un, pn)
        xc = x.detach().contiguous().to(torch.complex128)
        y = H.apply(xc)
        ctx.H, ctx.model, ctx.params = H, model, (tn, Un, pn)
        ctx.save_for_backward(xc)
        ctx.shapes = (t.shape, U.shape)
        return y

    @staticmethod
    def backward(ctx, gy):
        (xc,) = ctx.saved_tensors
        H, model = ctx.H, ctx.model
        tn, Un, pn = ctx.params
        gyc = gy.contiguous().to(torch.complex128)
        gx = H.apply(gyc)
        g_t, g_U, g_phi = model.grads(H, gyc, xc, tn, pn)
        dev = gy.device
        return (gx, torch.as_tensor(g_t, device=dev), torch.as_tensor(g_U, device=dev),
                torch.as_tensor(g_phi, dtype=torch.float64, device=dev), None)


def ground_state(model: ParamHubbard, t, U, phi, tol=1e-12, max_steps=1000):
    H = model.build(t, U, phi)
    f = torch.empty(H.dim, dtype=torch.complex128, device="cuda")
    ev = torch.empty_like(f)
    lz = Lanczos(H.dimUP, H.dimDN, max_steps=max_steps, tol=tol)
    lz.random_vector(f)
    e0 = lz.eigen(H, f, ev)
    return e0, ev, H


def ground_state_gradients(model: ParamHubbard, t, U, phi, tol=1e-12):
    """E0 and Hellmann-Feynman dE0/dt_b, dE0/dU_i, dE0/dphi."""
    e0, psi, H = ground_state(model, t, U, phi, tol)
    g_t, g_U, g_phi = model.grads(H, psi, psi, t, phi)
    return e0, g_t, g_U, g_phi


def kagome12_model(Qup=6, Qdn=6):
    """12-site Kagome (2x2 cells), L_x = 4 nearest-neighbour lengths (one x ring 0-1-3-4)."""
    return ParamHubbard(L.KAGOME12, Qup, Qdn, 4.0)

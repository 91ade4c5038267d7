"""Batched ground states and the twist-flux Drude-weight sweep.

BatchedLanczos solves many independent Hubbard instances together (ed_batch_*: one launch
per kernel for all instances, the convergence test on the device every ten steps, fifty-step
hipGraph replays); the plain path runs the same instances one after another through
Lanczos.eigen without a Ritz vector.  Both follow the
reference's Lanczos<T>::eigen (MBL_test/include/HubMatrix_GPU.cuh:561-613) per instance.

Drude weight from the flux curvature (Kohn): for every particle-number sector s of the
Lieb-12 cluster (TotalSpinExcited_Lieb_dilute.cu:147-159 bonds) the x rings carry a total
Peierls phase Phi_k = 2 pi k / K (spread evenly over the ring's bonds, see
lattices.lieb_twisted_ops), and

    D_s = [E0_s(Phi_1) - 2 E0_s(Phi_0) + E0_s(Phi_{K-1})] / (h^2 * N_sites),  h = 2 pi / K,

the central second difference at Phi = 0 (E0 is even and 2 pi periodic in Phi).  The same
formula is applied to the batched, plain and oracle E0 values.
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np
import torch

from ._lib import LanczosCfg, call, lib, stream_ptr
from .fermi_operator import H_TBq
from .hubmatrix import Hubbard, Lanczos
from . import lattices as L


class BatchedLanczos:
    def __init__(self, hubbards, complex_vectors: bool):
        self.hubbards = list(hubbards)
        arr = (C.c_void_p * len(self.hubbards))(*[h.handle.value for h in self.hubbards])
        b = C.c_void_p()
        call("ed_batch_create", len(self.hubbards), arr, int(complex_vectors), C.byref(b))
        self._b = b
        self.complex_vectors = complex_vectors

    def __del__(self):
        try:
            if getattr(self, "_b", None):
                call("ed_batch_destroy", self._b)
                self._b = None
        except Exception:
            pass

    def set_shape(self, dn_rows=4, dn_panels=4, up_rows=1, up_panels=4):
        """panel shape of the batch kernels (results do not depend on it)"""
        call("ed_batch_set_shape", self._b, dn_rows, dn_panels, up_rows, up_panels)

    def autotune(self, steps=30) -> dict:
        """time the panel shapes on a solve of `steps` steps and keep the fastest; returns the choice"""
        from ._lib import BatchTuning
        t = BatchTuning()
        call("ed_batch_autotune", self._b, int(steps), C.byref(t), stream_ptr())
        return {"dn_rows": t.dn_rows, "dn_panels": t.dn_panels, "up_rows": t.up_rows, "up_panels": t.up_panels,
                "seconds": t.seconds, "steps": t.steps, "candidate_s": [round(v, 4) for v in t.candidate_s]}

    def lowest(self, max_steps=1000, tol=None, use_graph=True, seed=235):
        """E0 of every instance (ed_batch_lowest).  use_graph: 0/False eager launches, 1/True
        fifty-step graph replays, 2 one graph launch for every step after the first ten up to
        max_steps.  Returns (evals [n], nsteps [n], status [n])."""
        n = len(self.hubbards)
        cfg = LanczosCfg(max_steps, 0.0 if tol is None else tol, 10, 0)
        ev = (C.c_double * n)()
        ns = (C.c_int * n)()
        st = (C.c_int * n)()
        call("ed_batch_lowest", self._b, C.byref(cfg), seed, int(use_graph), ev, ns, st, stream_ptr())
        return np.array(ev[:]), np.array(ns[:]), np.array(st[:])


    def lowest_k(self, nev: int, steps=600, tol=None, use_graph=True, seed=235, evecs=None):
        """The nev lowest eigenvalues of every instance from one batched run of at most `steps`
        steps, each instance stopped once its nev lowest pass (ed_batch_lowest_k: Cullum-Willoughby
        filter of each tridiagonal, ed_lanczos_eigen's residual test on each kept value).  evecs (optional): one device tensor per instance
        (its dim, float64 or complex128 as the batch) that gets the instance's normalised ground
        state when its status is 0 (a second batched run rebuilds the Ritz vectors).  Returns
        (evals [n, nev], nsteps [n], status [n])."""
        n = len(self.hubbards)
        cfg = LanczosCfg(int(steps), 0.0 if tol is None else tol, 10, 0)
        ev = np.zeros(n * nev)
        ns = (C.c_int * n)()
        st = (C.c_int * n)()
        vp = None
        if evecs is not None:
            if len(evecs) != n:
                raise ValueError("lowest_k: one evec per instance")
            want = torch.complex128 if self.complex_vectors else torch.float64
            for H, v in zip(self.hubbards, evecs):
                if not (v.is_cuda and v.dtype == want and v.is_contiguous() and v.numel() == H.dim):
                    raise ValueError(f"lowest_k: evecs must be contiguous {want} CUDA tensors of the instances' dims")
            vp = (C.c_void_p * n)(*[v.data_ptr() for v in evecs])
        call("ed_batch_lowest_k", self._b, C.byref(cfg), int(nev), seed, int(use_graph),
             ev.ctypes.data_as(C.POINTER(C.c_double)), ns, st, vp, stream_ptr())
        return ev.reshape(n, nev), np.array(ns[:]), np.array(st[:])


def plain_lowest(hubbards, complex_vectors: bool, max_steps=1000, tol=None, strict=False):
    """One instance after another: the reference's loop over samples/sectors.  strict: raise when
    an instance misses the convergence test (otherwise its last Ritz value is returned)."""
    evs, ns = [], []
    for H in hubbards:
        dt = torch.complex128 if complex_vectors else torch.float64
        f = torch.empty(H.dim, dtype=dt, device="cuda")
        lz = Lanczos(H.dimUP, H.dimDN, max_steps=max_steps, tol=tol)
        lz.random_vector(f)
        cfg = LanczosCfg(lz.max_steps, lz.tol, 10, 0)
        e = C.c_double()
        n = C.c_int()
        st = lib().ed_lanczos_eigen(H.handle, int(complex_vectors), C.byref(cfg), C.c_void_p(f.data_ptr()),
                                    C.byref(e), None, None, None, C.byref(n), stream_ptr())
        if st not in (0, 4) or (strict and st != 0):
            from ._lib import check
            check(st)
        evs.append(e.value)
        ns.append(n.value)
    return np.array(evs), np.array(ns)


def drude_from_flux(E0, nflux, n_sites):
    """E0[..., k] at Phi_k = 2 pi k / nflux -> central second difference at Phi = 0 / N_sites."""
    E0 = np.asarray(E0)
    h = 2.0 * math.pi / nflux
    return (E0[..., 1] - 2.0 * E0[..., 0] + E0[..., nflux - 1]) / (h * h * n_sites)


def lieb12_instances(U=4.0, nflux=16, t=1.0, n_total=12, sectors=None, fluxes=None):
    """(sector N_up, flux index k, Hubbard) for the Lieb-12 twist sweep, complex tables."""
    if sectors is None:
        sectors = list(range(n_total + 1))
    if fluxes is None:
        fluxes = list(range(nflux))
    out = []
    for k in fluxes:
        ops = L.lieb_twisted_ops(2.0 * math.pi * k / nflux, t=t)
        for nu in sectors:
            nd = n_total - nu
            hu = H_TBq(12, nu, complex_tables=True)
            hu.add_ops(ops)
            hd = H_TBq(12, nd, complex_tables=True)
            hd.add_ops(ops)
            out.append((nu, k, Hubbard(hu, hd, U)))
    return out


def lieb12_drude_sweep(U=4.0, nflux=16, batched=True, use_graph=1, max_steps=1000, tol=None, shape=None,
                       graph_steps=300):
    """All 13 sectors x nflux fluxes.  Returns (E0[13, nflux], D[13], seconds, nsteps).  shape: the
    batch kernels' panel shape (dn_rows, dn_panels, up_rows, up_panels), e.g. an autotuned one.
    max_steps: the step cap of both loops, the reference's maxN = 1000 (HubMatrix_GPU.cuh:563).
    use_graph 1 (default): fifty-step graph replays; 2: the batched solve after its first ten
    steps as one graph launch, capped at graph_steps (the instances converge in at most 130 steps).
    An instance that does not converge raises, in either loop."""
    inst = lieb12_instances(U, nflux)
    hs = [h for (_, _, h) in inst]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if batched:
        bl = BatchedLanczos(hs, True)
        if shape is not None:
            bl.set_shape(*shape)
        cap = min(max_steps, graph_steps) if use_graph == 2 else max_steps
        ev, ns, st = bl.lowest(cap, tol, use_graph)
        if np.any(st != 0):
            raise RuntimeError(f"lieb12_drude_sweep: {int(np.sum(st != 0))} instances did not converge in {cap} steps")
    else:
        ev, ns = plain_lowest(hs, True, max_steps, tol, strict=True)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    E0 = np.zeros((13, nflux))
    for (nu, k, _), e in zip(inst, ev):
        E0[nu, k] = e
    return E0, drude_from_flux(E0, nflux, 12), secs, ns

"""Disorder ensembles of the open Hubbard chain (the MBL drivers' model,
MBL_test/src/Hubbard_EE_test.cu:50-140: Line16 bonds with matrix element -t and an on-site
potential Make_N(i, eps_i) per species).

For each seeded realisation eps_i ~ Uniform[-W, W] (the same potential for both species;
numpy PCG64 seeded with seed0 + r) this module computes
  * E0 of every realisation in one batched Lanczos launch (ed_batch_lowest),
  * the nev lowest eigenvalues and the ground state (the cuSpectra solver role): by default from
    one batched Lanczos run of all realisations and a batched rebuild of their ground states
    (ed_batch_lowest_k, method "batched"), or one Lanczos run per realisation with selective
    orthogonalisation (ed_lanczos_lowest, method "lanczos"), or the thick-restart solver
    (ed_eigs_lowest, method "restart"),
  * the half-chain entanglement entropy S_vN of the ground state from the particle-number blocks
    of rho_A -- for all realisations at once, each block size solved for every state in one
    batched call (ed_entanglement_entropy_vn_batch) -- and optionally the reference's
    S_ref = -sum sigma ln sigma (per-block SVDs, ed_entanglement_entropy, per realisation),
  * the Hellmann-Feynman gradient dE0/d eps_i = <n_{i up}> + <n_{i dn}> (ed_site_density).
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np
import torch

from ._lib import call, dptr, stream_ptr
from .batched import BatchedLanczos
from .fermi_operator import H_TBq
from .hubmatrix import Hubbard, Lanczos
from .entangle import entropy_vn_batch
from .spectra import HubbardOp, SymEigsSolver
from . import lattices as L


def chain_disorder_ops(L_sites: int, eps, t: float = 1.0):
    """Line16 (Hubbard_EE_test.cu:119-140) with uniform hopping: Make_H(i, i+1, -t), then
    Make_N(i, eps_i)."""
    ops = [(L.OP_HOP, i, i + 1, -float(t)) for i in range(L_sites - 1)]
    ops += [(L.OP_DENSITY, i, i, float(eps[i])) for i in range(L_sites)]
    return ops


def realisation_eps(r: int, L_sites: int, W: float, seed0: int = 2024):
    return np.random.default_rng(seed0 + r).uniform(-W, W, L_sites)


def build_realisation(eps, L_sites=14, n_up=7, n_dn=7, U=4.0, t=1.0) -> Hubbard:
    ops = chain_disorder_ops(L_sites, eps, t)
    hu = H_TBq(L_sites, n_up)
    hu.add_ops(ops)
    hd = H_TBq(L_sites, n_dn)
    hd.add_ops(ops)
    return Hubbard(hu, hd, U)


@dataclass
class EnsembleResult:
    eps: np.ndarray          # [R, L]
    e0_batched: np.ndarray   # [R]
    evals: np.ndarray        # [R, nev]
    entropy: np.ndarray      # [R, 2] (S_vN, S_ref)
    grad_eps: np.ndarray     # [R, L]
    seconds: float
    seconds_batched_e0: float
    fallbacks: int = 0       # method "batched": realisations whose batched run did not pass the test

    @property
    def realisations_per_s(self):
        return len(self.eps) / self.seconds


def site_density(H: Hubbard, psi: torch.Tensor) -> np.ndarray:
    """[2, N]: <psi|n_{i sigma}|psi> (psi normalised)."""
    out = np.zeros(2 * H.N)
    call("ed_site_density", H.handle, int(psi.is_complex()), dptr(psi), dptr(psi),
         out.ctypes.data_as(C.POINTER(C.c_double)), stream_ptr())
    return out.reshape(2, H.N)


def half_chain_entropy(H: Hubbard, n_up: int, n_dn: int, psi: torch.Tensor, LA: int, with_sref=True):
    """(S_vN, S_ref) of sites 0..LA-1; S_ref (per-block SVDs) is nan when with_sref is False."""
    out = np.full(2, np.nan)
    if with_sref:
        nsv = C.c_int64(0)
        call("ed_entanglement_entropy", H.N, n_up, n_dn, (1 << LA) - 1, int(psi.is_complex()), dptr(psi),
             out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(nsv), stream_ptr())
    else:
        call("ed_entanglement_entropy_vn", H.N, n_up, n_dn, (1 << LA) - 1, int(psi.is_complex()), dptr(psi),
             out.ctypes.data_as(C.POINTER(C.c_double)), stream_ptr())
    return out


def run_ensemble(n_real=128, L_sites=14, n_up=7, n_dn=7, U=4.0, t=1.0, W=8.0, seed0=2024, nev=4, ncv=16,
                 tol=1e-10, maxit=1000, log=None, with_sref=True, workers=1, method="batched",
                 max_steps=2000, batch_steps=600, batch_size=128) -> EnsembleResult:
    """workers > 1 solves that many realisations at once after the batched launch, each worker
    thread on its own stream.  method "batched": the nev lowest of every realisation and its
    ground state from one batched run of at most batch_steps steps (each realisation stops once
    its nev lowest pass) and a batched rebuild of the Ritz vectors (ed_batch_lowest_k, test
    tolerance tol); a realisation whose batched values fail the test falls back to "lanczos".
    "lanczos": after a batched E0 launch, ed_lanczos_lowest per realisation (at most max_steps
    steps; tol is its test's tolerance, |b s_N| < tol |T| for each of the nev lowest Ritz pairs);
    "restart": the thick-restart solver with ncv basis vectors and maxit restarts (Spectra's test
    with tol).  Without S_ref the entropies are taken after the solves, all realisations of a
    batch in one batched call.  Realisations go through in batches of at most batch_size, so that
    the device memory held (the batch's vectors and ground states: about 6 batch_size vectors of
    the sector) does not grow with n_real."""
    if method not in ("batched", "lanczos", "restart"):
        raise ValueError("run_ensemble: method is 'batched', 'lanczos' or 'restart'")
    if batch_size < 1:
        raise ValueError("run_ensemble: batch_size >= 1")
    eps = np.stack([realisation_eps(r, L_sites, W, seed0) for r in range(n_real)])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tb = 0.0
    fallbacks = 0
    e0b = np.zeros(n_real)
    evals = np.zeros((n_real, nev))
    S = np.zeros((n_real, 2))
    g = np.zeros((n_real, L_sites))
    hs, psib, psis, evb, stb = {}, {}, {}, {}, {}
    import threading
    lock = threading.Lock()

    def solve(r, stream):
        # the library releases the GIL in every call; each worker thread drives its own stream
        nonlocal fallbacks
        with torch.cuda.stream(stream):
            H = hs[r]
            if method == "restart":
                es = SymEigsSolver(HubbardOp(H), nev, ncv)
                es.init()
                es.compute("SmallestAlge", maxit, tol)
                if es.info() != "Successful":
                    raise RuntimeError(f"realisation {r}: eigensolver did not converge")
                evals[r] = es.eigenvalues()
                psi = es.eigenvectors()[0]
                del es
            elif method == "batched" and stb[r] == 0:
                evals[r] = evb[r]
                psi = psib.pop(r)  # rebuilt by the batched call
            else:
                if method == "batched":
                    with lock:
                        fallbacks += 1
                    psib.pop(r, None)
                f = torch.empty(H.dim, dtype=torch.float64, device="cuda")
                psi = torch.empty_like(f)
                lz = Lanczos(H.dimUP, H.dimDN, max_steps=max_steps, tol=tol)
                lz.random_vector(f)
                evals[r] = lz.lowest(H, f, nev, psi)
                del f
            if with_sref:
                S[r] = half_chain_entropy(H, n_up, n_dn, psi, L_sites // 2, True)
            else:
                psis[r] = psi
            g[r] = site_density(H, psi).sum(0)
            stream.synchronize()
            del psi
            if log is not None:
                log(f"realisation {r}: E0 {evals[r, 0]:.12f}, {time.perf_counter() - t0:.2f} s elapsed")

    for c0 in range(0, n_real, batch_size):
        idx = list(range(c0, min(n_real, c0 + batch_size)))
        for r in idx:
            hs[r] = build_realisation(eps[r], L_sites, n_up, n_dn, U, t)
        hl = [hs[r] for r in idx]
        tc = time.perf_counter()
        if method == "batched":
            vecs = [torch.empty(h.dim, dtype=torch.float64, device="cuda") for h in hl]
            ev, _, st = BatchedLanczos(hl, False).lowest_k(nev, steps=batch_steps, tol=tol, evecs=vecs)
            for q, r in enumerate(idx):
                psib[r], evb[r], stb[r], e0b[r] = vecs[q], ev[q], st[q], ev[q, 0]
            del vecs
        else:
            e0b[idx], _, _ = BatchedLanczos(hl, False).lowest(tol=1e-12)
        torch.cuda.synchronize()
        tb += time.perf_counter() - tc
        if workers <= 1:
            for r in idx:
                solve(r, torch.cuda.current_stream())
        else:
            from concurrent.futures import ThreadPoolExecutor
            local = threading.local()

            def task(r):
                if not hasattr(local, "stream"):
                    local.stream = torch.cuda.Stream()  # one stream per worker thread
                solve(r, local.stream)

            # the batch's first realisation runs on this thread first: every library routine the
            # workers call (and the rocBLAS/rocSOLVER kernels behind them) is then loaded before
            # several threads make their first calls at once
            task(idx[0])
            with ThreadPoolExecutor(workers) as ex:
                for f in [ex.submit(task, r) for r in idx[1:]]:
                    f.result()
        if not with_sref:
            S[idx, 0] = entropy_vn_batch(L_sites, n_up, n_dn, (1 << (L_sites // 2)) - 1, [psis[r] for r in idx])
            S[idx, 1] = np.nan
            psis.clear()
        hs.clear()
        psib.clear()
    torch.cuda.synchronize()
    return EnsembleResult(eps, e0b, evals, S, g, time.perf_counter() - t0, tb, fallbacks)

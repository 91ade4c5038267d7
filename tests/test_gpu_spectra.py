"""Several lowest eigenpairs (ed_eigs_lowest, the cuSpectra::SymEigsSolver role), the host-array
perform_op, and the disorder-ensemble pipeline (lowest eigenvalues, half-chain entropy, dE0/d eps_i)
against the CPU oracle: dense numpy eigendecompositions of the oracle's H on small sectors, ARPACK
(scipy eigsh) over the oracle's H.v on larger ones, and the committed 14-site chain fixture
(tests/golden/chain14_eigs_oracle.json, written by tests/golden/make_chain14_eigs.py).

Bars: eigenvalues 1e-10 relative; entropies and site densities 1e-8 (ground states converged to a
residual of 1e-13 |E0|; the 8-site W = 8 realisations have gaps down to 1e-3); perform_op within
1e-14 of max|y| (same kernels as H.v).
"""
import json
import math
import os

import numpy as np
import pytest
import torch
from scipy.sparse.linalg import LinearOperator, eigsh

from oracle import oracle as O
from oracle import ee_oracle as EE
from exact_diagonalization_for_fermionic_systems_with_gpu_amd import lattices as L
from exact_diagonalization_for_fermionic_systems_with_gpu_amd import H_TBq, Hubbard, Lanczos
from exact_diagonalization_for_fermionic_systems_with_gpu_amd import disorder as DS
from exact_diagonalization_for_fermionic_systems_with_gpu_amd.spectra import HubbardOp, SymEigsSolver

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "chain14_eigs_oracle.json")


def _pair(N, Qu, Qd, ops, U, cplx=False):
    hu = H_TBq(N, Qu, complex_tables=cplx)
    hu.add_ops(ops)
    hd = H_TBq(N, Qd, complex_tables=cplx)
    hd.add_ops(ops)
    return Hubbard(hu, hd, U), O.Hubbard(O.Species(N, Qu, ops, complex_tables=cplx),
                                         O.Species(N, Qd, ops, complex_tables=cplx), U)


def _dense(Ho):
    D = Ho.dim
    e = np.zeros(D, dtype=Ho.dtype())
    M = np.zeros((D, D), dtype=Ho.dtype())
    for k in range(D):
        e[:] = 0
        e[k] = 1
        M[:, k] = Ho.apply(e)
    return M


def test_eigs_and_perform_op_vs_dense_chain8():
    eps = DS.realisation_eps(0, 8, 8.0)
    ops = DS.chain_disorder_ops(8, eps)
    Hg, Ho = _pair(8, 4, 4, ops, 4.0)
    print("eigs chain8 4/4 W=8", flush=True)
    w, V = np.linalg.eigh(_dense(Ho))
    es = SymEigsSolver(HubbardOp(Hg), 4, 16)
    es.init()
    assert es.compute("SmallestAlge", 1000, 1e-13) == 4 and es.info() == "Successful"
    ev = es.eigenvalues()
    assert np.allclose(ev, w[:4], rtol=1e-10, atol=0), (ev, w[:4])
    X = es.eigenvectors()
    for k in range(4):
        y = Hg.apply(X[k])
        r = float((y - ev[k] * X[k]).norm())
        assert r < 1e-9 * abs(ev[k]), (k, r)
    # perform_op on host arrays
    x = np.random.default_rng(1).normal(size=Hg.dim)
    y = np.empty_like(x)
    HubbardOp(Hg).perform_op(x, y)
    yo = Ho.apply(x)
    assert np.abs(y - yo).max() <= 1e-14 * np.abs(yo).max()


def test_eigs_complex_twist_lieb12_vs_arpack():
    ops = L.lieb_twisted_ops(2 * math.pi * 3 / 16, t=1.0)
    Hg, Ho = _pair(12, 3, 9, ops, 4.0, cplx=True)
    print("eigs lieb12 3/9 twisted (complex)", flush=True)
    A = LinearOperator((Ho.dim, Ho.dim), matvec=lambda x: Ho.apply(np.ascontiguousarray(x).ravel()),
                       dtype=np.complex128)
    wa = np.sort(eigsh(A, k=5, which="SA", tol=1e-13, ncv=30)[0])
    es = SymEigsSolver(HubbardOp(Hg, True), 5, 24)
    es.init()
    assert es.compute("SmallestAlge", 1000, 1e-12) == 5
    assert np.allclose(es.eigenvalues(), wa, rtol=1e-10, atol=1e-12), (es.eigenvalues(), wa)


def test_eigs_kagome12_vs_lanczos_and_arpack():
    ops = L.hubbard_ops(L.KAGOME12, t=-1.0)
    Hg, Ho = _pair(12, 6, 6, ops, 4.0)
    print("eigs kagome12 6/6", flush=True)
    es = SymEigsSolver(HubbardOp(Hg), 4, 20)
    es.init()
    assert es.compute("SmallestAlge", 1000, 1e-12) == 4
    ev = es.eigenvalues()
    f = torch.empty(Hg.dim, dtype=torch.float64, device="cuda")
    psi = torch.empty_like(f)
    lz = Lanczos(Hg.dimUP, Hg.dimDN, tol=1e-12)
    lz.random_vector(f)
    e0 = lz.eigen(Hg, f, psi)
    assert abs(ev[0] - e0) < 1e-10 * abs(e0)
    A = LinearOperator((Ho.dim, Ho.dim), matvec=lambda x: Ho.apply(np.ascontiguousarray(x).ravel()),
                       dtype=np.float64)
    wa = np.sort(eigsh(A, k=4, which="SA", tol=1e-13, ncv=24, v0=O.random_vector(Ho.dim, 235))[0])
    assert np.allclose(ev, wa, rtol=1e-10, atol=0), (ev, wa)


@pytest.mark.parametrize("method", ["batched", "lanczos", "restart"])
def test_disorder_ensemble_vs_dense_chain8(method):
    print(f"disorder ensemble chain8 x4 ({method})", flush=True)
    # the Lanczos test scales with |T| (about 3 |E0| here), the restart solver's with |theta|: S_ref
    # needs the ground state to ~1e-11, so the Lanczos run gets the tighter bar
    tol = 1e-14 if method == "lanczos" else 1e-13
    res = DS.run_ensemble(n_real=4, L_sites=8, n_up=4, n_dn=4, W=8.0, nev=4, ncv=16, tol=tol, method=method)
    # the batched run goes 400 steps on D = 4900, far past convergence: its tridiagonals hold
    # copies of the converged eigenvalues, which the Cullum-Willoughby filter has to remove
    assert res.fallbacks == 0
    for r in range(4):
        ops = DS.chain_disorder_ops(8, res.eps[r])
        Ho = O.Hubbard(O.Species(8, 4, ops), O.Species(8, 4, ops), 4.0)
        w, V = np.linalg.eigh(_dense(Ho))
        assert np.allclose(res.evals[r], w[:4], rtol=1e-10, atol=0)
        assert abs(res.e0_batched[r] - w[0]) < 1e-10 * abs(w[0])
        psi = V[:, 0]
        lam, sig = EE.block_spectra(8, 4, 4, 0b1111, psi)
        assert abs(res.entropy[r, 0] - EE.entropy_from_lambda(lam)) < 1e-8
        assert abs(res.entropy[r, 1] - EE.entropy_from_sigma(sig)) < 1e-8
        up, dn = Ho.up, Ho.dn
        p2 = psi.reshape(up.count, dn.count) ** 2
        nu = np.array([[(int(s) >> i) & 1 for i in range(8)] for s in up.state], float)
        nd = np.array([[(int(s) >> i) & 1 for i in range(8)] for s in dn.state], float)
        assert np.allclose(res.grad_eps[r], p2.sum(1) @ nu + p2.sum(0) @ nd, atol=1e-8)
        assert abs(res.grad_eps[r].sum() - 8.0) < 1e-10  # particle number


def test_disorder_gradient_vs_finite_differences():
    eps = DS.realisation_eps(3, 8, 8.0)
    h = 1e-5

    def e0(e):
        H = DS.build_realisation(e, 8, 4, 4)
        es = SymEigsSolver(HubbardOp(H), 2, 12)
        es.compute("SmallestAlge", 1000, 1e-13)
        return es.eigenvalues()[0], es.eigenvectors()[0], H

    E, psi, H = e0(eps)
    g = DS.site_density(H, psi).sum(0)
    for i in (0, 3, 7):
        ep, em = eps.copy(), eps.copy()
        ep[i] += h
        em[i] -= h
        fd = (e0(ep)[0] - e0(em)[0]) / (2 * h)
        assert abs(fd - g[i]) < 1e-6, (i, fd, g[i])


@pytest.mark.skipif(not os.path.exists(GOLD), reason="fixture not generated")
@pytest.mark.parametrize("method", ["batched", "lanczos"])
def test_chain14_against_fixture(method):
    gold = json.load(open(GOLD))
    print(f"disorder chain14 7/7 x2 vs fixture ({method})", flush=True)
    # S_ref to 1e-8 needs the ground state to ~1e-10: the Lanczos test at 1e-14 |T|; the batched
    # run's test at 1e-12 |T| (its 500 steps meet it)
    tol = 1e-12 if method == "batched" else 1e-14
    res = DS.run_ensemble(n_real=2, L_sites=14, n_up=7, n_dn=7, W=8.0, nev=4, ncv=20, tol=tol, method=method,
                          batch_steps=500)
    if method == "batched":
        assert res.fallbacks == 0
    for rec in gold["realisations"]:
        r = rec["r"]
        assert np.allclose(res.eps[r], rec["eps"], rtol=0, atol=0)
        assert np.allclose(res.evals[r], rec["evals"], rtol=1e-10, atol=0), (res.evals[r], rec["evals"])
        assert abs(res.e0_batched[r] - rec["evals"][0]) < 1e-10 * abs(rec["evals"][0])
        assert abs(res.entropy[r, 0] - rec["S_vN"]) < 1e-8
        assert abs(res.entropy[r, 1] - rec["S_ref"]) < 1e-8
        assert np.allclose(res.grad_eps[r], rec["dE0_deps"], atol=1e-8)


@pytest.mark.parametrize("method", ["batched", "lanczos"])
def test_disorder_ensemble_in_batches_matches_one_batch(method):
    """Realisations taken in batches of 4 (two batched launches) give what one batch of all 6
    gives, to the rounding of the other group size in the batched runs (the batch's partial sums
    are per tile of 256 / G columns)."""
    kw = dict(n_real=6, L_sites=8, n_up=4, n_dn=4, W=8.0, nev=3, tol=1e-12, method=method, with_sref=False)
    a = DS.run_ensemble(**kw)
    b = DS.run_ensemble(**kw, batch_size=4)
    assert np.allclose(a.evals, b.evals, rtol=1e-12, atol=0)
    assert np.allclose(a.entropy[:, 0], b.entropy[:, 0], rtol=0, atol=1e-9)
    assert np.allclose(a.grad_eps, b.grad_eps, rtol=0, atol=1e-9)
    assert np.allclose(a.e0_batched, b.e0_batched, rtol=1e-12, atol=0)


@pytest.mark.parametrize("method", ["lanczos", "batched"])
def test_disorder_ensemble_concurrent_workers_match_serial(method):
    """Several realisations solved (method "lanczos") or post-processed (entropies, densities) at
    once on separate streams (one rocBLAS handle and scratch per host thread) give the serial
    results."""
    kw = dict(n_real=6, L_sites=8, n_up=4, n_dn=4, W=8.0, nev=3, ncv=12, tol=1e-12, method=method)
    a = DS.run_ensemble(**kw, workers=1)
    b = DS.run_ensemble(**kw, workers=3)
    assert np.allclose(a.evals, b.evals, rtol=1e-12, atol=0)
    assert np.allclose(a.entropy, b.entropy, rtol=0, atol=1e-10)
    assert np.allclose(a.grad_eps, b.grad_eps, rtol=0, atol=1e-10)


@pytest.mark.parametrize("case", ["kagome12_u4", "chain8_disorder", "chain10_twist"])
def test_lanczos_lowest_vs_dense(case):
    """ed_lanczos_lowest: the nev lowest eigenvalues from one recurrence with selective
    orthogonalisation against the dense spectrum of the oracle's H (no spurious copies of
    converged eigenvalues), and its ground-state vector against the dense one."""
    print(f"lanczos lowest {case}", flush=True)
    cplx = False
    if case == "kagome12_u4":
        # 12-site Kagome at 5/5 with random hopping r_i in [0.8, 1.2] (no spatial symmetry, so no
        # exactly degenerate levels within the sector; dim 792^2 = 627k)
        r = np.random.default_rng(5).uniform(0.8, 1.2, 12)
        ops, N, nu, nd, U = L.hubbard_ops(L.KAGOME12, r=r), 12, 5, 5, 4.0
    elif case == "chain8_disorder":
        ops, N, nu, nd, U = DS.chain_disorder_ops(8, DS.realisation_eps(7, 8, 8.0)), 8, 4, 4, 4.0
    else:
        # periodic 8-site ring with a twist of 0.7 on the closing bond (complex128) and a random
        # on-site potential (no translation symmetry, so no degenerate levels within the sector)
        eps = np.random.default_rng(9).uniform(-1.0, 1.0, 8)
        ops = [(L.OP_HOP, i, i + 1, -1.0) for i in range(7)] + [(L.OP_HOP_PHASE, 7, 0, -1.0, 0.7)]
        ops += [(L.OP_DENSITY, i, i, float(eps[i])) for i in range(8)]
        N, nu, nd, U = 8, 3, 3, 2.0
        cplx = True
    hu = H_TBq(N, nu, complex_tables=cplx)
    hu.add_ops(ops)
    hd = H_TBq(N, nd, complex_tables=cplx)
    hd.add_ops(ops)
    Hg = Hubbard(hu, hd, U)
    Ho = O.Hubbard(O.Species(N, nu, ops, complex_tables=cplx), O.Species(N, nd, ops, complex_tables=cplx), U)
    nev = 4
    dt = torch.complex128 if cplx else torch.float64
    f = torch.empty(Hg.dim, dtype=dt, device="cuda")
    psi = torch.empty_like(f)
    lz = Lanczos(Hg.dimUP, Hg.dimDN, max_steps=3000, tol=1e-12)
    lz.random_vector(f)
    ev = lz.lowest(Hg, f, nev, psi)
    if Ho.dim <= 6000:
        w, V = np.linalg.eigh(_dense(Ho))
        ref = w[:nev]
        v0 = V[:, 0]
    else:
        # too large for a dense reference: the thick-restart solver (itself checked against ARPACK
        # over the oracle's H.v in test_eigs_kagome12_vs_lanczos_and_arpack) at a tight tolerance
        es = SymEigsSolver(HubbardOp(Hg), nev, 24)
        es.init()
        es.compute("SmallestAlge", 1000, 1e-13)
        assert es.info() == "Successful"
        ref = es.eigenvalues()
        v0 = es.eigenvectors()[0].cpu().numpy()
    assert np.allclose(ev, ref, rtol=1e-10, atol=0), (ev, ref)
    ov = abs(np.vdot(v0, psi.cpu().numpy()))
    assert abs(ov - 1.0) < 1e-8, ov
    assert lz.nsteps < 3000


def test_entropy_vn_batch_matches_per_state():
    """S_vN of many states at once (one batched eigenvalue call per block size) against the
    per-state entry point, real and complex, on the 7|7 cut of the 14-site chain (64 blocks up to
    1225 rows) and a 5|5 cut of a 10-site sector."""
    from exact_diagonalization_for_fermionic_systems_with_gpu_amd import entangle as EN
    print("entropy vn batch", flush=True)
    for (N, nu, nd, cplx, ns) in [(14, 7, 7, False, 5), (10, 5, 4, True, 3)]:
        dim = math.comb(N, nu) * math.comb(N, nd)
        g = torch.Generator(device="cpu").manual_seed(11 + N)
        states = []
        for _ in range(ns):
            if cplx:
                x = torch.randn(dim, 2, generator=g, dtype=torch.float64)
                x = torch.view_as_complex(x).contiguous()
            else:
                x = torch.randn(dim, generator=g, dtype=torch.float64)
            x = x / torch.linalg.vector_norm(x)
            states.append(x.cuda())
        mask = (1 << (N // 2)) - 1
        sb = EN.entropy_vn_batch(N, nu, nd, mask, states)
        for i, x in enumerate(states):
            out = np.zeros(1)
            import ctypes as C
            from exact_diagonalization_for_fermionic_systems_with_gpu_amd._lib import call, dptr, stream_ptr
            call("ed_entanglement_entropy_vn", N, nu, nd, mask, int(cplx), dptr(x),
                 out.ctypes.data_as(C.POINTER(C.c_double)), stream_ptr())
            assert abs(sb[i] - out[0]) < 1e-12 * max(1.0, abs(out[0])), (N, i, sb[i], out[0])
        # the oracle on one state
        lam, _ = EE.block_spectra(N, nu, nd, mask, states[0].cpu().numpy())
        assert abs(sb[0] - EE.entropy_from_lambda(lam)) < 1e-10


def test_batched_lowest_k_against_per_instance():
    """ed_batch_lowest_k on the 14-site chain (D = 1.18e7, one group of 6 realisations sharing the
    operator structure): the 4 lowest of every realisation against ed_lanczos_lowest on each, and
    the batched ground states (the rebuild run) against the plain Lanczos run's Ritz vector from
    the same start vector (ed_lanczos_eigen, same test), with their residuals."""
    from exact_diagonalization_for_fermionic_systems_with_gpu_amd.batched import BatchedLanczos
    print("batched lowest_k chain14 x6 with ground states", flush=True)
    hs = [DS.build_realisation(DS.realisation_eps(r, 14, 8.0), 14, 7, 7) for r in range(6)]
    psib = [torch.full((H.dim,), float("nan"), dtype=torch.float64, device="cuda") for H in hs]
    ev, ns, st = BatchedLanczos(hs, False).lowest_k(4, steps=500, tol=1e-10, evecs=psib)
    # the device test stops each realisation once its 4 lowest pass (240-330 steps here); one whose
    # lowest levels lie close together may need more than 500 (ED_ENOCONV, the ensemble then
    # solves it alone); the others must agree with the per-instance solver
    assert (ns <= 500).all() and (st == 0).sum() >= 4 and (ns[st == 0] < 500).all(), (st, ns)
    for r, H in enumerate(hs):
        if st[r] != 0:
            continue
        f = torch.empty(H.dim, dtype=torch.float64, device="cuda")
        lz = Lanczos(H.dimUP, H.dimDN, max_steps=2000, tol=1e-10)
        lz.random_vector(f)
        ref = lz.lowest(H, f, 4)
        assert np.allclose(ev[r], ref, rtol=1e-10, atol=0), (r, ev[r], ref)
        psi = torch.empty_like(f)
        lz.random_vector(f)
        e0 = lz.eigen(H, f, psi)
        pb = psib[r]
        assert abs(float(torch.linalg.vector_norm(pb)) - 1.0) < 1e-12
        if ev[r][1] - ev[r][0] > 1e-6:  # within a near-degenerate pair the vector is any mixture
            assert abs(abs(float(torch.dot(pb, psi))) - 1.0) < 1e-9, (r, float(torch.dot(pb, psi)))
        res = float(torch.linalg.vector_norm(H.apply(pb) - ev[r][0] * pb))
        assert res < 1e-7 and abs(e0 - ev[r][0]) < 1e-10 * abs(e0), (r, res, e0, ev[r][0])


def test_lanczos_lowest_near_degenerate_pairs_chain14():
    """Realisation 25 of the 14-site W = 8 chain has two pairs of levels 1e-8 apart
    (-46.81513937, -46.81513936, -46.00713629, -46.00713628): selective orthogonalisation counts a
    Ritz vector good long before it resolves such a pair, and without the Cullum-Willoughby filter
    the solver reported a third copy of the lowest pair.  Against the thick-restart solver."""
    print("lanczos lowest chain14 realisation 25 (near-degenerate pairs)", flush=True)
    H = DS.build_realisation(DS.realisation_eps(25, 14, 8.0), 14, 7, 7)
    f = torch.empty(H.dim, dtype=torch.float64, device="cuda")
    lz = Lanczos(H.dimUP, H.dimDN, max_steps=3000, tol=1e-10)
    lz.random_vector(f)
    ev = lz.lowest(H, f, 4)
    es = SymEigsSolver(HubbardOp(H), 4, 24)
    es.init()
    es.compute("SmallestAlge", 1000, 1e-12)
    assert es.info() == "Successful"
    ref = es.eigenvalues()
    assert ref[1] - ref[0] < 1e-7 and ref[3] - ref[2] < 1e-7  # the two close pairs
    assert np.allclose(ev, ref, rtol=1e-10, atol=0), (ev, ref)


def test_batched_lowest_k_near_degenerate_pairs_chain14():
    """realisation 25 (two pairs of levels 1e-8 apart) by the batched solve, against the
    thick-restart solver"""
    from exact_diagonalization_for_fermionic_systems_with_gpu_amd.batched import BatchedLanczos
    H = DS.build_realisation(DS.realisation_eps(25, 14, 8.0), 14, 7, 7)
    es = SymEigsSolver(HubbardOp(H), 4, 24)
    es.init()
    es.compute("SmallestAlge", 1000, 1e-12)
    ref = es.eigenvalues()
    evb, _, st = BatchedLanczos([H], False).lowest_k(4, steps=500, tol=1e-10)
    assert st[0] == 0 and np.allclose(evb[0], ref, rtol=1e-10, atol=0), (evb, ref)


def test_lanczos_lowest_start_vector_barely_touching_ground_state_chain14():
    """Realisation 125 of the 14-site W = 8 chain: the seed-235 start vector overlaps its ground
    state by 1.1e-7, so the converged lowest Ritz value is also an eigenvalue of T without its
    first row and column; the spurious-value filter once dropped it and the solver returned the
    first excited level (-11.7151) as E0.  Against the plain Lanczos E0."""
    print("lanczos lowest chain14 realisation 125 (start vector 1e-7 from the ground state)", flush=True)
    H = DS.build_realisation(DS.realisation_eps(125, 14, 8.0), 14, 7, 7)
    f = torch.empty(H.dim, dtype=torch.float64, device="cuda")
    psi = torch.empty_like(f)
    lz = Lanczos(H.dimUP, H.dimDN, max_steps=3000, tol=1e-12)
    lz.random_vector(f)
    f0 = f.clone()
    e0 = lz.eigen(H, f, psi)
    assert abs(float(torch.dot(f0, psi))) < 1e-6  # the case: a start vector nearly orthogonal to psi0
    f.copy_(f0)
    lw = Lanczos(H.dimUP, H.dimDN, max_steps=2000, tol=1e-10)
    ev = lw.lowest(H, f, 4, psi)
    assert abs(ev[0] - e0) < 1e-10 * abs(e0), (ev, e0)


def test_batched_lowest_k_start_vector_barely_touching_ground_state_chain14():
    """realisation 125 (start vector 1.1e-7 from the ground state) by the batched solve"""
    from exact_diagonalization_for_fermionic_systems_with_gpu_amd.batched import BatchedLanczos
    H = DS.build_realisation(DS.realisation_eps(125, 14, 8.0), 14, 7, 7)
    f = torch.empty(H.dim, dtype=torch.float64, device="cuda")
    psi = torch.empty_like(f)
    lz = Lanczos(H.dimUP, H.dimDN, max_steps=3000, tol=1e-12)
    lz.random_vector(f)
    e0 = lz.eigen(H, f, psi)
    evb, _, st = BatchedLanczos([H], False).lowest_k(4, steps=400, tol=1e-10)
    assert st[0] == 0 and abs(evb[0][0] - e0) < 1e-10 * abs(e0), (evb, e0)

"""GPU Krylov real-time evolution (Krylov<C>, HubMatrix_GPU.cuh:619-728) against the oracle's
restatement and the exact propagator.  Bars: eigenpairs of the tridiagonal 1e-12 (relative to
|T|); the first 20 Lanczos coefficients 1e-11 relative to the oracle's; propagated states
within 1e-9 of the exact exp(-i H dt) phi0 (chain) and of the oracle (12-site Kagome); norm
kept to 1e-9; the 6|6 entropy of the evolved state within 1e-8 of the oracle's block spectra."""
import numpy as np
import pytest
import torch

from oracle import ee_oracle as E
from oracle import oracle as O
from exact_diagonalization_for_fermionic_systems_with_gpu_amd import H_TBq, Hubbard, Krylov, tridiag_eigen
from exact_diagonalization_for_fermionic_systems_with_gpu_amd import lattices as L
from exact_diagonalization_for_fermionic_systems_with_gpu_amd.entangle import EEmatHubAB_GPU

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _chain_ops(Ls, eps):
    ops = [(L.OP_HOP, i, i + 1, -1.0) for i in range(Ls - 1)]
    return ops + [(L.OP_DENSITY, i, i, float(eps[i])) for i in range(Ls)]


def _pair(N, Qu, Qd, ops, U):
    hu, hd = H_TBq(N, Qu), H_TBq(N, Qd)
    hu.add_ops(ops)
    hd.add_ops(ops)
    Ho = O.Hubbard(O.Species(N, Qu, ops, True), O.Species(N, Qd, ops, True), U)
    return Hubbard(hu, hd, U), Ho


@pytest.mark.parametrize("n,seed,gap", [(300, 1, 1e-1), (200, 2, 1e-9)])
def test_tridiag_eigen_vs_lapack(n, seed, gap):
    print(f"tridiag_eigen n={n} seed={seed}", flush=True)
    r = np.random.default_rng(seed)
    a = r.normal(size=n)
    b = np.concatenate([[0.0], r.uniform(0.05, 1.5, n - 1)])
    if gap < 1e-3:  # two weakly coupled identical blocks: pairs of eigenvalues ~gap apart
        h = n // 2
        a[h:] = a[:h]
        b[h + 1:] = b[1:h]
        b[h] = gap
    m, z = tridiag_eigen(a, b)
    mo, zo = O.tridiag_eigen(a, b)
    T = np.diag(a) + np.diag(b[1:n], 1) + np.diag(b[1:n], -1)
    nT = np.abs(T).sum(1).max()
    assert np.abs(m - mo).max() <= 1e-12 * nT
    Z = z.reshape(n, n)
    assert np.abs(Z @ Z.T - np.eye(n)).max() < 1e-10
    assert np.abs(T @ Z.T - Z.T * m).max() < 1e-12 * nT


def test_krylov_chain8_vs_exact_and_oracle():
    Ls = 8
    eps = np.random.default_rng(2024).uniform(-8, 8, Ls)
    H, Ho = _pair(Ls, 4, 4, _chain_ops(Ls, eps), 4.0)
    print(f"krylov chain8 D={H.dim}", flush=True)
    phi0 = np.zeros(H.dim, dtype=np.complex128)
    phi0[H.dim // 3] = 1.0
    g0 = torch.from_numpy(phi0).to(DEV)
    kr = Krylov(H.dimUP, H.dimDN)
    kr.Make_tridiagonal_eigen(H, g0)
    ao, bo = O.krylov_tridiagonal(Ho, phi0, 40, 1e-10)
    assert np.allclose(kr.a[:20], ao[:20], rtol=1e-11, atol=1e-11)
    assert np.allclose(kr.b[1:21], bo[1:21], rtol=1e-11, atol=0)
    D = H.dim
    M = np.zeros((D, D), dtype=np.complex128)
    e = np.zeros(D, dtype=np.complex128)
    for k in range(D):
        e[:] = 0
        e[k] = 1
        M[:, k] = Ho.apply(e)
    w, V = np.linalg.eigh(M)
    for dt in (0.02, 0.5):
        phi = torch.zeros_like(g0)
        kr.Krylov_time_evol(H, g0, phi, dt)
        got = phi.cpu().numpy()
        exact = V @ (np.exp(-1j * w * dt) * (V.conj().T @ phi0))
        assert abs(np.linalg.norm(got) - 1.0) < 1e-9, dt
        assert np.abs(got - exact).max() < 1e-9, dt


def test_krylov_kagome12_quench_vs_oracle():
    ops = L.hubbard_ops(L.KAGOME12, t=-1.0)
    H, Ho = _pair(12, 6, 6, ops, 4.0)
    print(f"krylov kagome12 D={H.dim}", flush=True)
    hu = H_TBq(12, 6)
    up = int(hu.address[0b000000111111].item())
    dn = int(hu.address[0b111111000000].item())
    phi0 = np.zeros(H.dim, dtype=np.complex128)
    phi0[up * H.dimDN + dn] = 1.0
    g0 = torch.from_numpy(phi0).to(DEV)
    kr = Krylov(H.dimUP, H.dimDN, max_steps=80)
    kr.Make_tridiagonal_eigen(H, g0)
    assert len(kr.a) == 80
    ao, bo = O.krylov_tridiagonal(Ho, phi0, 80, 1e-10)
    mo, zo = O.tridiag_eigen(ao, bo)
    dt = 0.2
    phi = torch.zeros_like(g0)
    kr.Krylov_time_evol(H, g0, phi, dt)
    got = phi.cpu().numpy()
    ref = O.krylov_time_evol(Ho, phi0, ao, bo, mo, zo, dt)
    assert abs(np.linalg.norm(got) - 1.0) < 1e-9
    assert np.abs(got - ref).max() < 1e-9
    ee = EEmatHubAB_GPU(list(range(6)), 6, hu, hu)
    svn, _, _ = ee.entropy(phi)
    lam, _ = E.block_spectra(12, 6, 6, (1 << 6) - 1, ref)
    assert abs(svn - E.entropy_from_lambda(lam)) < 1e-8

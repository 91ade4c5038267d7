"""Pins of the oracle's Krylov real-time evolution (Krylov<C>, HubMatrix_GPU.cuh:619-728), CPU only.

* tridiag_eigen (LAPACK through scipy) against numpy's dense eigh of the same tridiagonal;
* when the Krylov space of phi0 is exhausted (b < 1e-10), Krylov_time_evol equals the exact
  propagator exp(-i H dt) phi0 from a dense eigendecomposition (a sign slip in the phase, a
  dropped b term or a transposed z fails it);
* on a larger sector with a truncated space the state stays normalised and matches the exact
  propagator for a short step.
"""
import numpy as np
import scipy.linalg as sla

from oracle import oracle as O
from exact_diagonalization_for_fermionic_systems_with_gpu_amd import lattices as L


def _ops(Ls, eps):
    ops = [(L.OP_HOP, i, i + 1, -1.0) for i in range(Ls - 1)]
    return ops + [(L.OP_DENSITY, i, i, float(eps[i])) for i in range(Ls)]


def _dense(H):
    D = H.dim
    M = np.zeros((D, D), dtype=np.complex128)
    e = np.zeros(D, dtype=np.complex128)
    for k in range(D):
        e[:] = 0
        e[k] = 1
        M[:, k] = H.apply(e)
    return M


def test_tridiag_eigen_matches_dense_eigh():
    r = np.random.default_rng(7)
    n = 60
    a = r.normal(size=n)
    b = np.concatenate([[0.0], r.uniform(0.1, 2.0, n - 1)])
    m, z = O.tridiag_eigen(a, b)
    T = np.diag(a) + np.diag(b[1:n], 1) + np.diag(b[1:n], -1)
    w, V = np.linalg.eigh(T)
    assert np.allclose(m, w, rtol=0, atol=1e-12)
    Z = z.reshape(n, n)  # row j = eigenvector j
    for j in range(n):
        assert np.linalg.norm(T @ Z[j] - m[j] * Z[j]) < 1e-12
        assert abs(abs(Z[j] @ V[:, j]) - 1.0) < 1e-10


def test_krylov_exhausted_space_is_exact_propagator():
    Ls = 4
    eps = np.random.default_rng(3).uniform(-2, 2, Ls)
    ops = _ops(Ls, eps)
    up, dn = O.Species(Ls, 1, ops, True), O.Species(Ls, 1, ops, True)
    H = O.Hubbard(up, dn, 2.0)
    M = _dense(H)
    phi0 = np.zeros(H.dim, dtype=np.complex128)
    phi0[5] = 1.0
    a, b = O.krylov_tridiagonal(H, phi0, 1000, 1e-10)
    assert len(a) <= H.dim and b[-1] < 1e-10
    m, z = O.tridiag_eigen(a, b)
    for dt in (0.0, 0.3, 1.7):
        phi = O.krylov_time_evol(H, phi0, a, b, m, z, dt)
        exact = sla.expm(-1j * dt * M) @ phi0
        assert np.abs(phi - exact).max() < 1e-10, dt


def test_krylov_truncated_space_short_step():
    Ls = 7
    eps = np.random.default_rng(11).uniform(-8, 8, Ls)
    ops = _ops(Ls, eps)
    up, dn = O.Species(Ls, 3, ops, True), O.Species(Ls, 3, ops, True)
    H = O.Hubbard(up, dn, 4.0)
    M = _dense(H)
    phi0 = np.zeros(H.dim, dtype=np.complex128)
    phi0[617] = 1.0
    a, b = O.krylov_tridiagonal(H, phi0, 60, 1e-10)
    assert len(a) == 60
    m, z = O.tridiag_eigen(a, b)
    dt = 0.05
    phi = O.krylov_time_evol(H, phi0, a, b, m, z, dt)
    w, V = np.linalg.eigh(M)
    exact = V @ (np.exp(-1j * w * dt) * (V.conj().T @ phi0))
    assert abs(np.linalg.norm(phi) - 1.0) < 1e-10
    assert np.abs(phi - exact).max() < 1e-10

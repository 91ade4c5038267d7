"""The spurious-value filter behind ed_lanczos_lowest and ed_batch_lowest_k
(ed_tridiag_lanczos_lowest, host code), on Lanczos tridiagonals made here in numpy from diagonal
matrices with known spectra:

* a start vector that touches the ground state by 1e-8 (full reorthogonalisation, so T has no
  ghosts): the converged lowest Ritz value is then also an eigenvalue of T without its first row
  and column (checked), and the plain Cullum-Willoughby test would call it spurious.  After
  selective orthogonalisation the filter keeps it; after a plain recurrence it holds the test
  back (until a copy forms).  A disordered 14-site chain had a start vector 1.1e-7 from its
  ground state, and the solver reported its first excited level as E0;
* a long run without reorthogonalisation (ghost copies of the low end present, checked): every
  check that reports convergence reports the nev lowest eigenvalues of the matrix, each once;
* sweeps over start vectors that barely touch one of the four lowest eigenvalues (1e-12 to 1e-6),
  with and without reorthogonalisation: no check reports convergence with a wrong set once every
  one of the four has a Ritz value near it.
"""
import numpy as np
import pytest
import scipy.linalg as sla
import ctypes as C

from exact_diagonalization_for_fermionic_systems_with_gpu_amd import _lib


def lanczos(lam, f, m, full_reorth):
    """m steps of the Lanczos recurrence on diag(lam) from f: a[m], b[m + 1] (b[0] = 0, b[m] the
    coupling after the last step)"""
    n = lam.size
    a = np.zeros(m)
    b = np.zeros(m + 1)
    V = np.zeros((m, n)) if full_reorth else None
    v_prev = np.zeros(n)
    v = f / np.linalg.norm(f)
    for k in range(m):
        if full_reorth:
            V[k] = v
        w = lam * v - b[k] * v_prev
        a[k] = v @ w
        w -= a[k] * v
        if full_reorth:
            for _ in range(2):
                w -= V[: k + 1].T @ (V[: k + 1] @ w)
        b[k + 1] = np.linalg.norm(w)
        v_prev, v = v, w / b[k + 1]
    return a, b


def lowest(a, b, nev, tol, semi_orth):
    m = a.size
    anorm = max(abs(a[0] + b[1]), max(b[k] + abs(a[k] + b[k + 1]) for k in range(1, m)))
    out = (C.c_double * nev)()
    found = C.c_int()
    conv = C.c_int()
    st = _lib.lib().ed_tridiag_lanczos_lowest(m, a.ctypes.data_as(_lib._pd), b[1:m].copy().ctypes.data_as(_lib._pd),
                                              float(b[m]), nev, tol, anorm, int(semi_orth), out, C.byref(found),
                                              C.byref(conv))
    assert st == 0
    return np.array(out[:]), found.value, bool(conv.value), anorm


@pytest.mark.parametrize("semi_orth", [0, 1])
def test_ground_state_barely_touched_by_start_vector_is_kept(semi_orth):
    rng = np.random.default_rng(7)
    n = 3000
    lam = np.concatenate([[-1.0, -0.93, -0.9], np.sort(rng.uniform(-0.85, 1.0, n - 3))])
    f = rng.standard_normal(n)
    f /= np.linalg.norm(f)
    f[0] = 1e-8
    a, b = lanczos(lam, f, 260, True)
    w, Z = sla.eigh_tridiagonal(a, b[1:-1])
    wh = sla.eigvalsh_tridiagonal(a[1:], b[2:-1])
    anorm = max(abs(a).max() + 2 * b.max(), 1.0)
    # the situation the filter has to get right: the lowest Ritz value has converged to -1 and
    # sits on an eigenvalue of T^
    assert abs(w[0] + 1.0) < 1e-12
    assert abs(Z[0, 0]) < 1e-6
    assert np.abs(wh - w[0]).min() < 1e-12 * anorm
    out, found, conv, _ = lowest(a, b, 3, 1e-10, semi_orth)
    assert found == 3 and conv == bool(semi_orth)  # a plain recurrence waits for its copy
    assert np.allclose(out, lam[:3], rtol=0, atol=1e-11), out


def test_ghost_copies_are_dropped_after_plain_recurrence():
    rng = np.random.default_rng(11)
    n = 1200
    lam = np.concatenate([[-2.0, -1.7, -1.5, -1.4], np.sort(rng.uniform(-1.2, 1.0, n - 4))])
    f = rng.standard_normal(n)
    a, b = lanczos(lam, f, 600, False)
    w = sla.eigvalsh_tridiagonal(a, b[1:-1])
    # loss of orthogonality has made copies of the lowest eigenvalue
    assert (np.abs(w - lam[0]) < 1e-9).sum() >= 2
    reports = 0
    for m in range(300, 601, 10):
        out, found, conv, _ = lowest(a[:m], b[:m + 1], 4, 1e-10, 0)
        if conv:
            reports += 1
            assert np.allclose(out, lam[:4], rtol=0, atol=1e-10), (m, out)
    assert reports >= 5


@pytest.mark.parametrize("reorth", [False, True])
def test_barely_touched_levels_never_give_a_wrong_set(reorth):
    n = 1200
    reports = 0
    for trial in range(6):
        rng = np.random.default_rng(300 + trial)
        low = np.sort(rng.uniform(-1.0, -0.85, 5))
        lam = np.concatenate([low, np.sort(rng.uniform(-0.8, 1.0, n - 5))])
        f = rng.standard_normal(n)
        f /= np.linalg.norm(f)
        f[trial % 4] = 10.0 ** rng.uniform(-12, -6)
        a, b = lanczos(lam, f, 300, reorth)
        for m in range(20, 301, 10):
            out, found, conv, _ = lowest(a[:m], b[:m + 1], 4, 1e-10, int(reorth))
            if not conv:
                continue
            w = sla.eigvalsh_tridiagonal(a[:m], b[1:m])
            if all(np.abs(w - lam[q]).min() < 1e-4 for q in range(4)):
                reports += 1
                assert np.allclose(out, lam[:4], rtol=0, atol=1e-9), (trial, m, out - lam[:4])
    assert reports >= 20


def test_unconverged_lowest_holds_the_test_back():
    # after a few steps nothing has converged: the filter reports no convergence in both modes
    rng = np.random.default_rng(3)
    n = 2000
    lam = np.sort(rng.uniform(-1.0, 1.0, n))
    a, b = lanczos(lam, rng.standard_normal(n), 12, True)
    for semi_orth in (0, 1):
        _, _, conv, _ = lowest(a, b, 2, 1e-10, semi_orth)
        assert not conv

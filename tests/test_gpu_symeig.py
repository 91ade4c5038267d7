"""Two-stage symmetric eigenvalue routine (ed_sym_eigvals: dense -> band by blocked Householder
panels, band -> tridiagonal by bulge chasing, Sturm bisection), the eigenvalue stage of the large
rho_A blocks, against LAPACK (numpy eigvalsh) on the same matrices.

Bar: every eigenvalue within 1e-12 |A|_2 (backward-stable reductions give a few eps |A| per
eigenvalue; LAPACK's own error is of the same order).  Orders cover the panel width 64 and its
neighbours, orders below it (no panel: the bulge chasing alone), 1-3, and one of several panels
with a ragged last one; the matrices include a Gram matrix with eigenvalues spread over 16 decades
(the rho_A case) and exactly repeated eigenvalues."""
import ctypes as C

import numpy as np
import pytest
import torch

from exact_diagonalization_for_fermionic_systems_with_gpu_amd._lib import call, stream_ptr

pytestmark = pytest.mark.gpu


def _gpu_eigvals(A, upper=False):
    n = A.shape[0]
    M = np.asfortranarray(A)
    if upper:
        M = np.asfortranarray(np.triu(A))
    d = torch.from_numpy(np.ascontiguousarray(M.T).reshape(-1)).cuda()  # column-major buffer
    w = torch.zeros(n, dtype=torch.float64, device="cuda")
    call("ed_sym_eigvals", n, C.c_void_p(d.data_ptr()), n, int(upper), C.c_void_p(w.data_ptr()), stream_ptr())
    return w.cpu().numpy()


def _cases(n, rng):
    G = rng.standard_normal((n, n))
    yield "symmetric", (G + G.T) / 2
    # Gram matrix with a spectrum over 16 decades, like a block of rho_A
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = 10.0 ** rng.uniform(-16, 0, n)
    yield "gram", (Q * lam) @ Q.T
    # exactly repeated eigenvalues
    lam = np.repeat([-1.0, 0.5, 2.0], (n + 2) // 3)[:n]
    yield "repeated", (Q * lam) @ Q.T


@pytest.mark.parametrize("n", [1, 2, 3, 5, 63, 64, 65, 127, 128, 129, 200, 515])
def test_sym_eigvals_vs_lapack(n):
    rng = np.random.default_rng(100 + n)
    print(f"sym eigvals n={n}", flush=True)
    for name, A in _cases(n, rng):
        ref = np.linalg.eigvalsh(A)
        scale = max(np.abs(ref).max(), 1e-300)
        for upper in (False, True):
            w = _gpu_eigvals(A, upper)
            assert np.all(np.diff(w) >= 0), (name, "not ascending")
            err = np.abs(w - ref).max() / scale
            assert err <= 1e-12, (name, upper, err)


def test_sym_eigvals_several_panels_entropy():
    """n = 2100 (32 full panels and a ragged one): eigenvalues and -sum lambda ln lambda of a
    unit-trace Gram matrix against LAPACK's."""
    n = 2100
    rng = np.random.default_rng(7)
    Psi = rng.standard_normal((n, n)) * np.exp(-np.arange(n) / 60.0)
    rho = Psi @ Psi.T
    rho /= np.trace(rho)
    print("sym eigvals n=2100", flush=True)
    w = _gpu_eigvals(rho, upper=True)
    ref = np.linalg.eigvalsh(rho)
    assert np.abs(w - ref).max() <= 1e-12 * ref.max()
    s = lambda x: -np.sum(x[x > 0] * np.log(x[x > 0]))
    assert abs(s(w) - s(ref)) < 1e-10

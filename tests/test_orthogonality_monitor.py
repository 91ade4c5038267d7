"""The omega recurrence (ed_lanczos_orthogonality, host only) against the true loss of
orthogonality of a Lanczos run without reorthogonalisation.

The run is the reference's recurrence (HubMatrix_GPU.cuh:573-582) written in numpy over the
oracle's H.v, with every Lanczos vector kept, so max_{k<j} |v_j . v_k| is measured directly.
The monitor follows Simon's model of the rounding, so it is an estimate, not a bound: the checks
are that it errs high (never below a third of the measured loss once that is above rounding
level) and that it reaches sqrt(eps) no later than the measured loss does, and at most 15 steps
earlier.  A dropped or mis-signed term of the recurrence breaks the growth rate and fails these.
"""
import numpy as np
import pytest

from exact_diagonalization_for_fermionic_systems_with_gpu_amd import lanczos_orthogonality
from exact_diagonalization_for_fermionic_systems_with_gpu_amd import lattices as L
from oracle import oracle as O


def plain_lanczos(N, nsteps, U=4.0, seed=1):
    ops = [(L.OP_HOP, i, (i + 1) % N, -1.0) for i in range(N)]
    sp = O.Species(N, N // 2, ops)
    H = O.Hubbard(sp, sp, U)
    f = np.random.default_rng(seed).standard_normal(H.dim)
    V = np.zeros((nsteps, H.dim))
    a, b, loss = np.zeros(nsteps), np.zeros(nsteps + 1), np.zeros(nsteps)
    b[0] = np.linalg.norm(f)
    v, vprev = f / b[0], np.zeros(H.dim)
    for j in range(nsteps):
        V[j] = v
        if j > 0:
            loss[j] = np.abs(V[:j] @ v).max()
        w = H.apply(v)
        if j > 0:
            w -= b[j] * vprev
        a[j] = v @ w
        w -= a[j] * v
        b[j + 1] = np.linalg.norm(w)
        vprev, v = v, w / b[j + 1]
    return a, b, loss


@pytest.mark.parametrize("N,nsteps", [(8, 110), (10, 120)])
def test_omega_estimate_tracks_measured_loss(N, nsteps):
    a, b, loss = plain_lanczos(N, nsteps)
    om = lanczos_orthogonality(a, b)
    se = np.sqrt(np.finfo(float).eps)
    assert om[0] == 0.0 and np.all(np.isfinite(om))
    cross_true = int(np.argmax(loss > se))
    cross_est = int(np.argmax(om > se))
    assert cross_true > 0, "the run must lose orthogonality"
    assert cross_true - 15 <= cross_est <= cross_true
    sel = (np.arange(nsteps) >= 10) & (loss > 1e-13) & (np.arange(nsteps) <= cross_true)
    assert np.all(om[sel] >= loss[sel] / 3.0)
    # deterministic
    assert np.array_equal(om, lanczos_orthogonality(a, b))


def test_omega_estimate_of_an_orthogonal_run_stays_at_rounding_level():
    """A short run (12 steps on the 6-site ring, before any Ritz value converges) keeps its basis
    orthogonal to rounding level, and the estimate says so."""
    a, b, loss = plain_lanczos(6, 12)
    om = lanczos_orthogonality(a, b)
    assert loss.max() < 1e-13
    assert om.max() < 1e-12

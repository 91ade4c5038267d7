"""Adjoint of y = H(theta) x and Hellmann-Feynman gradients (12-site Kagome, per-bond t_b,
per-site U_i, twist flux Phi, complex128).

* HubbardApply's gradients for x, t, U and Phi against torch autograd on a dense H(theta)
  assembled from the CPU oracle's per-call tables (a small cluster);
* dE0/dtheta (Hellmann-Feynman, Ritz vector rebuilt by recomputing Lanczos) against central
  finite differences of E0 with h = 1e-5 (tolerance 2e-6 relative: the FD truncation and the
  1e-13 accuracy of E0 divided by h);
* Euler's identity sum_b t_b dE0/dt_b + sum_i U_i dE0/dU_i = E0 (H is linear in t and U), to 1e-10;
* the flux Drude weight (1/N) d2E0/dPhi2 at Phi = 0 from the Hellmann-Feynman slope agrees with
  the second difference of E0.
"""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from exact_diagonalization_for_fermionic_systems_with_gpu_amd import adjoint as AD
from exact_diagonalization_for_fermionic_systems_with_gpu_amd import lattices as L

pytestmark = pytest.mark.gpu


def _params(seed=2024):
    r = np.random.default_rng(seed)
    return r.uniform(0.8, 1.2, 24), r.uniform(2.0, 6.0, 12), 0.3


def _ring6():
    bonds = [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (5, 0), (0, 3)]
    dx = [1.0, 0.5, -0.5, 1.0, 0.5, -0.5, 0.0]
    return L.Cluster("ring6", 6, bonds, [(i, j, 0.0, d) for (i, j), d in zip(bonds, dx)], [])


def _unit_parts(N, Q, op_ij):
    """P^f, P^b of the call for bond i -> j (c+_j c_i carries e^{+i phi}) on the Q sector,
    from the oracle's tables: M(phi) = e^{i phi} P^f + e^{-i phi} P^b."""
    i, j = op_ij

    def mat(phase):
        sp = O.Species(N, Q, [(L.OP_HOP_PHASE, j, i, 1.0, phase)], complex_tables=True)
        M = np.zeros((sp.count, sp.count), dtype=complex)
        for s in range(sp.slots):
            for a in range(sp.count):
                v = sp.Ax[s * sp.count + a]
                if v != 0:
                    M[a, sp.Aj[s * sp.count + a]] += v
        return M
    M0, M1 = mat(0.0), mat(math.pi / 2)
    return (M0 - 1j * M1) / 2, (M0 + 1j * M1) / 2


def test_autograd_matches_dense_torch():
    cl = _ring6()
    N, Qu, Qd, Lx = 6, 3, 2, 3.0
    model = AD.ParamHubbard(cl, Qu, Qd, Lx)
    rng = np.random.default_rng(0)
    t0, U0, phi0 = rng.uniform(0.8, 1.2, len(cl.bonds)), rng.uniform(1, 3, N), 0.4
    parts_u = [_unit_parts(N, Qu, (i, j)) for (i, j) in cl.bonds]
    parts_d = [_unit_parts(N, Qd, (i, j)) for (i, j) in cl.bonds]
    su = O.Species(N, Qu).state
    sd = O.Species(N, Qd).state
    docc = np.array([[[(int(a) & int(b)) >> k & 1 for b in sd] for a in su] for k in range(N)],
                    dtype=float).reshape(N, -1)
    du, dd = len(su), len(sd)
    # dense torch model
    t = torch.tensor(t0, requires_grad=True)
    U = torch.tensor(U0, requires_grad=True)
    phi = torch.tensor(phi0, dtype=torch.float64, requires_grad=True)
    dx = torch.tensor(model.dx)
    Hu = torch.zeros((du, du), dtype=torch.complex128)
    Hd = torch.zeros((dd, dd), dtype=torch.complex128)
    for k in range(len(cl.bonds)):
        ph = phi * dx[k] / Lx
        e = torch.exp(1j * ph)
        Hu = Hu + (-t[k]) * (e * torch.tensor(parts_u[k][0]) + torch.conj(e) * torch.tensor(parts_u[k][1]))
        Hd = Hd + (-t[k]) * (e * torch.tensor(parts_d[k][0]) + torch.conj(e) * torch.tensor(parts_d[k][1]))
    Hfull = torch.kron(Hu, torch.eye(dd, dtype=torch.complex128)) + torch.kron(torch.eye(du, dtype=torch.complex128),
                                                                             Hd)
    Hfull = Hfull + torch.diag((U[:, None] * torch.tensor(docc)).sum(0).to(torch.complex128))
    x0 = rng.normal(size=du * dd) + 1j * rng.normal(size=du * dd)
    w0 = rng.normal(size=du * dd) + 1j * rng.normal(size=du * dd)
    xd = torch.tensor(x0, requires_grad=True)
    y = Hfull @ xd
    loss = (torch.conj(torch.tensor(w0)) * y).real.sum() + (y.abs() ** 2).sum() * 0.1
    loss.backward()
    # GPU Function
    tg = torch.tensor(t0, requires_grad=True, device="cuda")
    Ug = torch.tensor(U0, requires_grad=True, device="cuda")
    pg = torch.tensor(phi0, dtype=torch.float64, requires_grad=True, device="cuda")
    xg = torch.tensor(x0, requires_grad=True, device="cuda")
    yg = AD.HubbardApply.apply(xg, tg, Ug, pg, model)
    assert torch.allclose(yg.detach().cpu(), y.detach(), atol=1e-12)
    lg = (torch.conj(torch.tensor(w0, device="cuda")) * yg).real.sum() + (yg.abs() ** 2).sum() * 0.1
    lg.backward()
    assert torch.allclose(xg.grad.cpu(), xd.grad, atol=1e-10)
    assert torch.allclose(tg.grad.cpu(), t.grad, atol=1e-9)
    assert torch.allclose(Ug.grad.cpu(), U.grad, atol=1e-9)
    assert abs(float(pg.grad) - float(phi.grad)) < 1e-9


def test_hellmann_feynman_vs_finite_differences_kagome12():
    model = AD.kagome12_model()
    t, U, phi = _params()
    print("HF gradients kagome12 6/6", flush=True)
    e0, g_t, g_U, g_phi = AD.ground_state_gradients(model, t, U, phi)
    # Euler: H is linear in (t, U)
    assert abs(np.dot(t, g_t) + np.dot(U, g_U) - e0) < 1e-10 * abs(e0)
    h = 1e-5
    for k in (0, 5, 13, 22):
        tp, tm = t.copy(), t.copy()
        tp[k] += h
        tm[k] -= h
        fd = (AD.ground_state(model, tp, U, phi)[0] - AD.ground_state(model, tm, U, phi)[0]) / (2 * h)
        assert abs(fd - g_t[k]) < 2e-6 * max(1.0, abs(g_t[k])), (k, fd, g_t[k])
    for i in (0, 7, 11):
        Up, Um = U.copy(), U.copy()
        Up[i] += h
        Um[i] -= h
        fd = (AD.ground_state(model, t, Up, phi)[0] - AD.ground_state(model, t, Um, phi)[0]) / (2 * h)
        assert abs(fd - g_U[i]) < 2e-6 * max(1.0, abs(g_U[i])), (i, fd, g_U[i])
    fd = (AD.ground_state(model, t, U, phi + h)[0] - AD.ground_state(model, t, U, phi - h)[0]) / (2 * h)
    assert abs(fd - g_phi) < 2e-6 * max(1.0, abs(g_phi)), (fd, g_phi)


def test_flux_drude_weight_from_slope():
    model = AD.kagome12_model()
    t, U, _ = _params()
    h = 1e-3
    gp = AD.ground_state_gradients(model, t, U, h)[3]
    gm = AD.ground_state_gradients(model, t, U, -h)[3]
    D_slope = (gp - gm) / (2 * h) / 12
    e = [AD.ground_state(model, t, U, p)[0] for p in (-h, 0.0, h)]
    D_curv = (e[0] - 2 * e[1] + e[2]) / (h * h) / 12
    # second differences of E0 lose ~1e-13/h^2 = 1e-7 absolute
    assert abs(D_slope - D_curv) < 5e-6 * max(1.0, abs(D_slope)), (D_slope, D_curv)
    assert abs(gp + gm) < 1e-8  # E0 even in Phi: the slope is odd

"""GPU parity of the dense full-spectrum path, the general fermion operators and the LOW_MEM
basis against the CPU oracle, through the C ABI of libedhip.so.

Bars: tables, dense writes and basis orders bit-exact (same adds in the same order per entry);
eigenvalues within 1e-12 of max|lambda| of LAPACK's (numpy eigh), eigenvector residuals
|H v - lambda v| <= 1e-11 |H|; tED within 1e-11 of scipy's matrix exponential (the propagation is
a sum of unit-modulus phases times unit vectors, rounding only).
"""
import math

import numpy as np
import pytest
import scipy.linalg
import torch

from oracle import oracle as O
from exact_diagonalization_for_fermionic_systems_with_gpu_amd import lattices as L
from exact_diagonalization_for_fermionic_systems_with_gpu_amd import (H_TBq, Hubbard, cusolverDsyevd, lowmem_basis,
                                                                       tED)
from exact_diagonalization_for_fermionic_systems_with_gpu_amd import totalspin

pytestmark = pytest.mark.gpu

DEV = "cuda"


def to_np(t):
    return t.detach().cpu().numpy()


FERMI_OPS = [(O.OP_FERMI1, 2, 2, 0.7, 0.0, 0), (O.OP_FERMI1, 4, 4, -1.3, 0.0, 1), (O.OP_FERMI2, 1, 4, 0.5, 0.0, 2),
             (O.OP_FERMI2, 5, 0, 1.1, 0.0, 1), (O.OP_FERMI2, 3, 1, 0.9, 0.0, 0), (O.OP_FERMI2, 0, 5, -0.4, 0.0, 3),
             (L.OP_HOP, 0, 1, -1.0), (O.OP_FERMI1, 5, 5, 0.25, 0.0, 0), (O.OP_FERMI2, 2, 3, 0.3, 0.0, 1),
             (L.OP_DENSITY, 3, 3, 0.5)]


def test_fermi_op_tables_bit_exact():
    print("fermi-op tables, full Fock N=6", flush=True)
    g = H_TBq(6)
    g.add_ops(FERMI_OPS)
    o = O.Species(6, -1, FERMI_OPS)
    assert np.array_equal(to_np(g.Ax).view(np.uint64), o.Ax.view(np.uint64))
    assert np.array_equal(to_np(g.Aj), o.Aj)


def test_fermi_op_methods_record_calls():
    g = H_TBq(5)
    g.Insert_1Fop("+", 2, 0.5)
    g.Insert_2Fop("-", 4, "+", 1, -0.75)
    assert g.ops == [(5, 2, 2, 0.5, 0.0, 0), (6, 4, 1, -0.75, 0.0, 1)]
    o = O.Species(5, -1, [(5, 2, 2, 0.5, 0.0, 0), (6, 4, 1, -0.75, 0.0, 1)])
    assert np.array_equal(to_np(g.Ax).view(np.uint64), o.Ax.view(np.uint64))
    with pytest.raises(ValueError):
        g.Insert_1Fop("x", 0, 1.0)


def test_full_fock_hubbard_with_fermi_ops_hv():
    print("H.v with Insert_1Fop/2Fop species (table kernels)", flush=True)
    up, dn = H_TBq(6), H_TBq(6)
    up.add_ops(FERMI_OPS)
    dn.add_ops(FERMI_OPS[::-1])
    H = Hubbard(up, dn, 2.0)
    OH = O.Hubbard(O.Species(6, -1, FERMI_OPS), O.Species(6, -1, FERMI_OPS[::-1]), 2.0)
    x = np.random.default_rng(1).normal(size=OH.dim)
    y = to_np(H.apply(torch.from_numpy(x).to(DEV)))
    yo = OH.apply(x)
    assert np.max(np.abs(y - yo)) <= 1e-14 * np.max(np.abs(yo))


def test_fermi_ops_rejected_by_on_the_fly_consumers():
    up, dn = H_TBq(6, 3), H_TBq(6, 3)
    up.add_ops([(L.OP_HOP, 0, 1, -1.0), (O.OP_FERMI2, 0, 2, 0.5, 0.0, 2)])
    dn.add_ops([(L.OP_HOP, 0, 1, -1.0)])
    H = Hubbard(up, dn, 1.0)
    from exact_diagonalization_for_fermionic_systems_with_gpu_amd._lib import EdError
    x = torch.ones(H.dim, dtype=torch.float64, device=DEV)
    with pytest.raises(EdError):
        H.autotune(x, torch.empty_like(x))


def _chain_species(L_, qu, qd, seed, cplx=False, W=2.0):
    rng = np.random.default_rng(seed)
    ops = L.hubbard_ops(L.chain(L_, periodic=True), t=-1.0)
    if cplx:
        ops = [(L.OP_HOP_PHASE, i, j, t, 0.3) for (_, i, j, t, *_) in ops]
    eu = [(L.OP_DENSITY, i, i, float(v)) for i, v in enumerate(rng.uniform(-W, W, L_))]
    ed = [(L.OP_DENSITY, i, i, float(v)) for i, v in enumerate(rng.uniform(-W, W, L_))]
    gu, gd = H_TBq(L_, qu, complex_tables=cplx), H_TBq(L_, qd, complex_tables=cplx)
    gu.add_ops(ops + eu)
    gd.add_ops(ops + ed)
    ou = O.Species(L_, qu, ops + eu, complex_tables=cplx)
    od = O.Species(L_, qd, ops + ed, complex_tables=cplx)
    return gu, gd, ou, od


@pytest.mark.parametrize("cplx", [False, True])
def test_hubbard_write_bit_exact(cplx):
    print(f"Hubbard::Write chain-8 4/3 cplx={cplx}", flush=True)
    gu, gd, ou, od = _chain_species(8, 4, 3, 3, cplx=cplx)
    H = Hubbard(gu, gd, 4.0)
    dt = torch.complex128 if cplx else torch.float64
    Hm = torch.empty((H.dim, H.dim), dtype=dt, device=DEV)
    H.initialize(Hm)
    H.Write(Hm)
    M = O.Hubbard(ou, od, 4.0).dense()
    assert np.array_equal(to_np(Hm), M)


def test_htbq_write_bit_exact():
    print("H_TBq::Write spinless chain-12 Q=6 and full Fock with Insert ops", flush=True)
    rng = np.random.default_rng(5)
    ops = L.hubbard_ops(L.chain(12), t=-1.0) + [(L.OP_DENSITY, i, i, float(v)) for i, v in
                                               enumerate(rng.uniform(-1, 1, 12))]
    g = H_TBq(12, 6)
    g.add_ops(ops)
    assert np.array_equal(to_np(g.Write()), O.Species(12, 6, ops).dense())
    g6 = H_TBq(6)
    g6.add_ops(FERMI_OPS)
    M = torch.full((64, 64), 0.5, dtype=torch.float64, device=DEV)  # Write adds into what is there
    g6.Write(M)
    assert np.array_equal(to_np(M), O.Species(6, -1, FERMI_OPS).dense() + 0.5)


@pytest.mark.parametrize("cplx", [False, True])
def test_dense_eigh(cplx):
    print(f"dense eigh chain-8 4/4 cplx={cplx}", flush=True)
    gu, gd, ou, od = _chain_species(8, 4, 4, 7, cplx=cplx)
    H = Hubbard(gu, gd, 4.0)
    dt = torch.complex128 if cplx else torch.float64
    A = torch.zeros((H.dim, H.dim), dtype=dt, device=DEV)
    H.Write(A)
    M = O.Hubbard(ou, od, 4.0).dense()
    W = torch.empty(H.dim, dtype=torch.float64, device=DEV)
    cusolverDsyevd(A, W)
    w, V = to_np(W), to_np(A)  # V[k] = eigenvector k
    wo = np.linalg.eigvalsh(M)
    scale = np.max(np.abs(wo))
    assert np.max(np.abs(w - wo)) <= 1e-12 * scale
    R = M @ V.T - V.T * w[None, :]
    assert np.max(np.abs(R)) <= 1e-11 * scale
    G = V.conj() @ V.T
    assert np.max(np.abs(G - np.eye(H.dim))) <= 1e-11


@pytest.mark.parametrize("cplx", [False, True])
def test_ted_time_evol_vs_expm(cplx):
    print(f"tED chain-8 4/4 cplx={cplx}", flush=True)
    gu, gd, ou, od = _chain_species(8, 4, 4, 11, cplx=cplx)
    H = Hubbard(gu, gd, 4.0)
    dt_ = torch.complex128 if cplx else torch.float64
    Hm = torch.zeros((H.dim, H.dim), dtype=dt_, device=DEV)
    H.Write(Hm)
    ted = tED(H.dimUP, H.dimDN)
    ted.Make_eigen(Hm)
    M = O.Hubbard(ou, od, 4.0).dense()
    phi0 = torch.zeros(H.dim, dtype=torch.complex128, device=DEV)
    k0 = int(gu.address[0b00001111].item()) * H.dimDN + int(gd.address[0b11110000].item())
    phi0[k0] = 1.0
    p0 = to_np(phi0)
    for t in (0.05, 1.0, 6.0):
        phi = torch.zeros_like(phi0)
        ted.tED_time_evol(phi0, phi, t)
        exact = scipy.linalg.expm(-1j * M * t) @ p0
        assert np.max(np.abs(to_np(phi) - exact)) <= 1e-11
        assert abs(np.linalg.norm(to_np(phi)) - 1.0) <= 1e-12
    # accumulation into a nonzero phi, as the reference's axpy chain does
    phi = phi0.clone()
    ted.tED_time_evol(phi0, phi, 0.0)
    assert np.max(np.abs(to_np(phi) - 2 * p0)) <= 1e-12


def test_total_spin_full_spectrum():
    print("TotalSpinFull flow: Write, syevd, S^2 of the lowest eigenvectors", flush=True)
    ops = L.hubbard_ops(L.chain(6, periodic=True), t=-1.0)
    gu, gd = H_TBq(6, 3), H_TBq(6, 3)
    gu.add_ops(ops)
    gd.add_ops(ops)
    H = Hubbard(gu, gd, 4.0)
    A = torch.zeros((H.dim, H.dim), dtype=torch.float64, device=DEV)
    H.Write(A)
    W = torch.empty(H.dim, dtype=torch.float64, device=DEV)
    cusolverDsyevd(A, W)
    S = totalspin.Hubbard(gu, gd)
    ou, od = O.Species(6, 3, ops), O.Species(6, 3, ops)
    V = to_np(A)
    for k in range(6):
        v = A[k].contiguous()
        sxy = float(torch.dot(v, S.Hubbard_mv_new(v)).item())
        sxy_o = float(V[k] @ O.s2_apply(ou, od, V[k]))
        assert abs(sxy - sxy_o) <= 1e-12
    # the half-filled 6-site ring has a nondegenerate singlet ground state: S(S+1) = Sxy = 0
    v0 = A[0].contiguous()
    assert abs(float(torch.dot(v0, S.Hubbard_mv_new(v0)).item())) <= 1e-10


@pytest.mark.parametrize("N,Q", [(4, 2), (10, 4), (16, 8), (40, 2), (63, 1)])
def test_lowmem_basis_bit_exact(N, Q):
    print(f"LOW_MEM basis N={N} Q={Q}", flush=True)
    g = to_np(lowmem_basis(N, Q)).view(np.uint64)
    if N <= 16:
        assert np.array_equal(g, O.lowmem_basis(N, Q))
    else:  # closed form: descending bit-reversed value
        rev = [int(format(int(s), f"0{N}b")[::-1], 2) for s in g]
        assert rev == sorted(rev, reverse=True) and len(g) == math.comb(N, Q)
        assert all(bin(int(s)).count("1") == Q for s in g)

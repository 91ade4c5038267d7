"""The C++ interface (include/ed/ed.hpp): tests/cpp/kagome12_driver.cpp is written against the
reference's class names (H_TBq, Hubbard, Lanczos<T>, TotalSpin::Hubbard, compute_sigma_omega,
EEmatHubAB_GPU, cuSpectra::Hubbard / SymEigsSolver, tED, cusolverDsyevd, H_TBq::Insert_1Fop /
Insert_2Fop / Write, LOW_MEM::H_TBq_basis) and built by build().  Its output for the 12-site Kagome
model (6/6, U = 4, hopping -1), an 8-site ring (dense path) and a 5-site Fock space is checked
against the CPU oracle.

Bars: E0, the cuSpectra eigenvalues and the host-array Rayleigh quotient 1e-10 relative; S(S+1)
1e-8; entropies 1e-8; the Kubo D_x 1e-7 relative (oracle run from its own ground state: the
GPU-vs-oracle difference from the same state is 6e-9, the rest is the ~1e-9 difference of the
two converged ground states amplified by the 200 unreorthogonalised Lanczos steps).
"""
import json
import os
import subprocess

import numpy as np
import pytest
from scipy.sparse.linalg import LinearOperator, eigsh

from oracle import oracle as O
from oracle import ee_oracle as EE
from exact_diagonalization_for_fermionic_systems_with_gpu_amd import lattices as L

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(__file__), "cpp", "bin", "kagome12_driver")


def test_kagome12_cpp_driver_against_oracle():
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} missing: run __graft_entry__.build()")
    print("C++ driver kagome12", flush=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    cl = L.KAGOME12
    ops = L.hubbard_ops(cl, t=-1.0)
    up, dn = O.Species(12, 6, ops), O.Species(12, 6, ops)
    H = O.Hubbard(up, dn, 4.0)
    st, e0, psi, a, b, n = H.lanczos_eigen(O.random_vector(H.dim, 235))
    assert st == 0
    assert abs(out["E0"] - e0) < 1e-10 * abs(e0)
    assert out["nsteps"] == n
    ss = 0.25 * 0 + float(psi @ O.s2_apply(up, dn, psi))
    assert abs(out["SS"] - ss) < 1e-8
    lam, sig = EE.block_spectra(12, 6, 6, 0b111111, psi)
    assert abs(out["S_vN"] - EE.entropy_from_lambda(lam)) < 1e-8
    assert abs(out["S_ref"] - EE.entropy_from_sigma(sig)) < 1e-8
    kx, _ = L.kinetic_ops(cl)
    jx, _ = L.current_ops(cl)
    J = O.Hubbard(O.Species(12, 6, jx), O.Species(12, 6, jx), 0.0)
    K = O.Hubbard(O.Species(12, 6, kx), O.Species(12, 6, kx), 0.0)
    D, kin, _ = O.drude_weight(H, J, K, e0, psi)
    assert abs(out["kinetic"] - kin) < 1e-10 * abs(kin)
    assert abs(out["D"] - D) < 1e-7 * abs(D), (out["D"], D)
    A = LinearOperator((H.dim, H.dim), matvec=lambda x: H.apply(np.ascontiguousarray(x).ravel()), dtype=np.float64)
    wa = np.sort(eigsh(A, k=3, which="SA", tol=1e-13, ncv=20, v0=O.random_vector(H.dim, 235))[0])
    assert out["nconv"] == 3
    assert np.allclose(out["evals"], wa, rtol=1e-10, atol=0)
    assert abs(out["rayleigh_host"] - e0) < 1e-10 * abs(e0)
    # Krylov quench of the product state (up on sites 0-5, dn on 6-11), dt = 0.2, 80 steps
    assert out["krylov_steps"] == 80
    assert abs(out["quench_norm2"] - 1.0) < 1e-9
    sq = O.Species(12, 6, ops, True)
    Hq = O.Hubbard(sq, sq, 4.0)
    phi0 = np.zeros(Hq.dim, dtype=np.complex128)
    adr = {int(s): k for k, s in enumerate(sq.state)}
    phi0[adr[0x03f] * Hq.dim_dn + adr[0xfc0]] = 1.0
    ak, bk = O.krylov_tridiagonal(Hq, phi0, 80, 1e-10)
    mk, zk = O.tridiag_eigen(ak, bk)
    phi = O.krylov_time_evol(Hq, phi0, ak, bk, mk, zk, 0.2)
    lamq, _ = EE.block_spectra(12, 6, 6, 0x3f, phi)
    assert abs(out["quench_S_vN"] - EE.entropy_from_lambda(lamq)) < 1e-8
    # the ground state with the recurrence on the slab layout
    assert abs(out["E0_slab"] - e0) < 1e-10 * abs(e0)
    assert out["nsteps_slab"] == n
    assert abs(abs(out["overlap_slab"]) - 1.0) < 1e-8
    # dense path on the 8-site ring, 4/4, U = 4: eigenvalues, tED amplitude and norm
    ring = [(L.OP_HOP, i, (i + 1) % 8, -1.0) for i in range(8)]
    Hr = O.Hubbard(O.Species(8, 4, ring), O.Species(8, 4, ring), 4.0)
    M = Hr.dense()
    w, V = np.linalg.eigh(M)
    assert np.allclose(out["ring_W"], w[:3], rtol=0, atol=1e-12 * np.abs(w).max())
    assert abs(out["ring_lanczos_E0"] - w[0]) < 1e-10 * abs(w[0])
    k0 = Hr.up.address[0x0f] * Hr.dim_dn + Hr.dn.address[0xf0]
    p0 = np.zeros(Hr.dim, dtype=np.complex128)
    p0[k0] = 1.0
    import scipy.linalg
    amp = (scipy.linalg.expm(-1j * M * 1.5) @ p0)[k0]
    assert abs(complex(*out["ted_amp0"]) - amp) < 1e-11
    assert abs(out["ted_norm2"] - 1.0) < 1e-12
    # H_TBq::Write with Insert_1Fop / Insert_2Fop on the 5-site Fock space
    F = O.Species(5, -1, [(O.OP_FERMI1, 2, 2, 0.5, 0.0, 0), (O.OP_FERMI2, 4, 1, -0.75, 0.0, 1),
                          (L.OP_DENSITY, 3, 3, 0.25)]).dense()
    assert out["fock5_checksum"] == float(np.sum((np.arange(32 * 32) + 1) * F.ravel()))
    assert out["lowmem_4_2"] == [3, 5, 9, 6, 10, 12]

"""Batched ground states (ed_batch_*) on the Lieb-12 twist sweep: against the plain one-by-one
Lanczos on the same GPU, against the CPU oracle's energies (tests/golden/lieb12_drude_oracle.json,
written by tests/golden/make_lieb12_drude.py), with and without hipGraph capture.  Bars: E0 within
1e-10 relative, Drude weights within 1e-8 relative (absolute 1e-12 where D = 0)."""
import json
import math
import os

import numpy as np
import pytest
import torch

from exact_diagonalization_for_fermionic_systems_with_gpu_amd import batched as BT

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "lieb12_drude_oracle.json")))


def _sweep_subset(use_graph, max_steps=1000):
    inst = BT.lieb12_instances(U=4.0, nflux=16, fluxes=[0, 1, 15])
    hs = [h for (_, _, h) in inst]
    ev, ns, st = BT.BatchedLanczos(hs, True).lowest(max_steps, use_graph=use_graph)
    return inst, hs, ev, ns, st


def test_batched_matches_plain_and_oracle():
    print("batched lieb12 subset", flush=True)
    inst, hs, ev, ns, st = _sweep_subset(True)
    assert np.all(st == 0)
    evp, nsp = BT.plain_lowest(hs, True)
    assert np.all(np.abs(ev - evp) <= 1e-10 * np.maximum(1.0, np.abs(evp)))
    E0 = np.zeros((13, 16))
    for (nu, k, _), e in zip(inst, ev):
        E0[nu, k] = e
    for nu in range(13):
        for k in (0, 1):
            g = GOLD["E0"][f"{nu}_{k}"]
            assert abs(E0[nu, k] - g) <= 1e-10 * max(1.0, abs(g)), (nu, k, E0[nu, k], g)
    D = BT.drude_from_flux(E0, 16, 12)
    for nu in range(13):
        g = GOLD["D"][nu]
        if abs(g) > 1e-6:
            assert abs(D[nu] - g) <= 1e-8 * abs(g), (nu, D[nu], g)
        else:
            assert abs(D[nu]) < 1e-12
    # Lieb's theorem: at half filling the ground state has S = |N_A - N_B|/2 = 2, so the
    # S_z = -2..2 sectors (N_up = 4..8) are degenerate
    for nu in range(4, 9):
        assert abs(E0[nu, 0] - E0[6, 0]) < 1e-10


def test_graph_replay_equals_eager():
    """fifty-step graph replays and eager launches: the same step counts and eigenvalues bit for
    bit"""
    _, _, ev_g, ns_g, _ = _sweep_subset(1)
    _, _, ev_e, ns_e, _ = _sweep_subset(0)
    assert np.array_equal(ns_g, ns_e)
    assert np.array_equal(ev_g.view(np.uint64), ev_e.view(np.uint64))


def test_one_graph_solve_equals_eager():
    """one graph for the whole solve after the first ten steps (capped at 300 steps) against eager
    launches: the same step counts and eigenvalues bit for bit"""
    _, _, ev_e, ns_e, _ = _sweep_subset(0)
    _, _, ev_w, ns_w, st_w = _sweep_subset(2, max_steps=300)
    assert np.array_equal(ns_w, ns_e) and np.all(st_w == 0)
    assert np.array_equal(ev_w.view(np.uint64), ev_e.view(np.uint64))


def test_full_sweep_all_instances():
    """All 208 instances of the headline sweep: batched E0 equals the plain per-instance E0, and
    E0 is even in the flux, E0(Phi_k) = E0(Phi_{16-k}), both to 1e-10 relative.  At Phi = pi the
    twisted complex operator gives the E0 of the reference's real pi-flux cluster (x bonds 3 and 7
    flipped, LiebLattice12 of TotalSpinExcited_Lieb_randomHop.cu:176-188) on the same GPU."""
    print("batched lieb12 full sweep", flush=True)
    E0b, Db, _, nsb = BT.lieb12_drude_sweep(batched=True)
    E0p, Dp, _, nsp = BT.lieb12_drude_sweep(batched=False)
    scale = np.maximum(1.0, np.abs(E0p))
    assert np.all(np.abs(E0b - E0p) <= 1e-10 * scale), np.max(np.abs(E0b - E0p))
    for E in (E0b, E0p):
        for k in range(1, 16):
            assert np.all(np.abs(E[:, k] - E[:, 16 - k]) <= 1e-10 * scale[:, k]), k
    from exact_diagonalization_for_fermionic_systems_with_gpu_amd import H_TBq, Hubbard
    from exact_diagonalization_for_fermionic_systems_with_gpu_amd import lattices as L
    ops = L.lieb_pi_flipped_ops()
    hs = []
    for nu in (3, 5, 6):
        hu, hd = H_TBq(12, nu), H_TBq(12, 12 - nu)
        hu.add_ops(ops)
        hd.add_ops(ops)
        hs.append(Hubbard(hu, hd, 4.0))
    epi, _ = BT.plain_lowest(hs, False)
    for e, nu in zip(epi, (3, 5, 6)):
        assert abs(E0b[nu, 8] - e) <= 1e-10 * max(1.0, abs(e)), (nu, E0b[nu, 8], e)


def test_batch_shapes_bitwise_and_autotune():
    """Every panel shape of the batch kernels gives the same energies bit for bit (per-element sums
    in call order, dot partials per 4-row unit), and the autotuner keeps one of them."""
    inst = BT.lieb12_instances(U=4.0, nflux=16, fluxes=[1, 5])
    hs = [h for (_, _, h) in inst]
    bl = BT.BatchedLanczos(hs, True)
    ref = None
    for shape in ((4, 4, 1, 4), (8, 1, 4, 16), (4, 8, 1, 1), (8, 2, 1, 8)):
        bl.set_shape(*shape)
        ev, ns, st = bl.lowest(use_graph=True)
        assert np.all(st == 0)
        if ref is None:
            ref = (ev, ns)
        else:
            assert np.array_equal(ns, ref[1]), shape
            assert np.array_equal(ev.view(np.uint64), ref[0].view(np.uint64)), shape
    t = bl.autotune(steps=20)
    assert t["dn_rows"] in (4, 8) and t["up_rows"] in (1, 4) and t["seconds"] > 0
    ev, ns, st = bl.lowest(use_graph=True)
    assert np.array_equal(ev.view(np.uint64), ref[0].view(np.uint64))


def test_lowest_k_complex_ground_states():
    """ed_batch_lowest_k with ground states on complex instances (Lieb-12 at flux 2 pi / 16, five
    sectors): values against the plain Lanczos E0, ground states against its Ritz vector from the
    same start vector, residual |H psi - E0 psi|.  With a 20-step cap nothing passes: every
    instance reports ED_ENOCONV and its vector is left as it was."""
    from exact_diagonalization_for_fermionic_systems_with_gpu_amd import Lanczos
    print("batched lowest_k complex lieb12 with ground states", flush=True)
    inst = BT.lieb12_instances(U=4.0, nflux=16, sectors=[3, 5, 6, 7, 9], fluxes=[1])
    hs = [h for (_, _, h) in inst]
    psib = [torch.full((h.dim,), complex("nan"), dtype=torch.complex128, device="cuda") for h in hs]
    ev, ns, st = BT.BatchedLanczos(hs, True).lowest_k(2, steps=600, tol=1e-10, evecs=psib)
    assert np.all(st == 0), (st, ns)
    for h, e, pb in zip(hs, ev, psib):
        f = torch.empty(h.dim, dtype=torch.complex128, device="cuda")
        psi = torch.empty_like(f)
        lz = Lanczos(h.dimUP, h.dimDN, max_steps=2000, tol=1e-10)
        lz.random_vector(f)
        e0 = lz.eigen(h, f, psi)
        assert abs(e[0] - e0) < 1e-10 * max(1.0, abs(e0)), (e, e0)
        assert abs(float(torch.linalg.vector_norm(pb)) - 1.0) < 1e-12
        if e[1] - e[0] > 1e-6:
            assert abs(abs(complex(torch.vdot(pb, psi))) - 1.0) < 1e-9
        res = float(torch.linalg.vector_norm(h.apply(pb) - e[0] * pb))
        assert res < 1e-7, res
    short = [torch.full((h.dim,), complex("nan"), dtype=torch.complex128, device="cuda") for h in hs]
    ev2, ns2, st2 = BT.BatchedLanczos(hs, True).lowest_k(2, steps=20, tol=1e-10, evecs=short)
    assert np.all(st2 != 0) and np.all(ns2 == 20), (st2, ns2)
    assert all(bool(torch.isnan(v.real).all()) for v in short)


def test_lowest_k_small_sectors_end_in_invariant_subspaces():
    """Sectors of dimension 1, 144 and 4356 (Lieb-12, N_up = 0, 1, 2 at flux 2 pi / 16): the two
    larger recurrences end in an invariant subspace before 300 steps or converge, their values
    and rebuilt ground states against the plain Lanczos solve; the dimension-1 sector has no
    second eigenvalue, so it reports ED_ENOCONV and keeps its vector."""
    from exact_diagonalization_for_fermionic_systems_with_gpu_amd import Lanczos
    print("batched lowest_k small sectors", flush=True)
    inst = BT.lieb12_instances(U=4.0, nflux=16, sectors=[0, 1, 2], fluxes=[1])
    hs = [h for (_, _, h) in inst]
    assert [h.dim for h in hs] == [1, 144, 4356]
    psib = [torch.full((h.dim,), complex("nan"), dtype=torch.complex128, device="cuda") for h in hs]
    ev, ns, st = BT.BatchedLanczos(hs, True).lowest_k(2, steps=300, tol=1e-10, evecs=psib)
    assert st[0] != 0 and bool(torch.isnan(psib[0].real).all())
    assert st[1] == 0 and st[2] == 0, (st, ns)
    for h, e, pb in zip(hs[1:], ev[1:], psib[1:]):
        f = torch.empty(h.dim, dtype=torch.complex128, device="cuda")
        psi = torch.empty_like(f)
        lz = Lanczos(h.dimUP, h.dimDN, max_steps=2000, tol=1e-10)
        lz.random_vector(f)
        e0 = lz.eigen(h, f, psi)
        assert abs(e[0] - e0) < 1e-10 * max(1.0, abs(e0)), (e, e0)
        res = float(torch.linalg.vector_norm(h.apply(pb) - e[0] * pb))
        assert abs(float(torch.linalg.vector_norm(pb)) - 1.0) < 1e-12 and res < 1e-7, res

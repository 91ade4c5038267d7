"""GPU parity of the slab layout (the Lanczos solver's internal vector layout): H.v on it gives
ed_hubbard_apply's values bit for bit (same terms, same per-element order), the layout
conversions are exact permutations with zero padding, and the Lanczos ground state on it meets
the E0 bar (1e-10 relative) with the same number of steps as on the reference layout.  Ragged
dimensions (not multiples of 16 or 32) exercise the padded slabs and partial tiles.
"""
import json
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O
from exact_diagonalization_for_fermionic_systems_with_gpu_amd import lattices as L
from exact_diagonalization_for_fermionic_systems_with_gpu_amd import H_TBq, Hubbard, Lanczos
from exact_diagonalization_for_fermionic_systems_with_gpu_amd._lib import EdError

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _model(name):
    rng = np.random.default_rng(17)
    if name == "chain8_4_3":
        ops = L.hubbard_ops(L.chain(8, periodic=True), t=-1.0)
        eu = [(L.OP_DENSITY, i, i, float(v)) for i, v in enumerate(rng.uniform(-2, 2, 8))]
        return 8, 4, 3, ops + eu, ops, 3.0, None
    if name == "kagome12_6_6":
        return 12, 6, 6, L.hubbard_ops(L.KAGOME12, t=-1.0), L.hubbard_ops(L.KAGOME12, t=-1.0), 4.0, None
    if name == "kagome12_randomhop_uvec":
        r = rng.uniform(0.8, 1.2, 12)
        ops = L.hubbard_ops(L.KAGOME12, r=r)
        return 12, 5, 7, ops, ops, 0.0, rng.uniform(2, 6, 12)
    if name in ("chain8_0_3", "chain8_4_8", "chain8_8_0", "chain8_3_3_dn_nocalls"):
        ops = L.hubbard_ops(L.chain(8, periodic=True), t=-1.0)
        eu = [(L.OP_DENSITY, i, i, float(v)) for i, v in enumerate(rng.uniform(-2, 2, 8))]
        qu, qd = {"chain8_0_3": (0, 3), "chain8_4_8": (4, 8), "chain8_8_0": (8, 0),
                  "chain8_3_3_dn_nocalls": (3, 3)}[name]
        return 8, qu, qd, ops + eu, ([] if name.endswith("nocalls") else ops), 2.5, None
    if name == "chain14_7_7":
        ops = L.hubbard_ops(L.chain(14), t=-1.0)
        ed = [(L.OP_DENSITY, i, i, float(v)) for i, v in enumerate(rng.uniform(-8, 8, 14))]
        return 14, 7, 7, ops + ed, ops + ed[::-1], 4.0, None
    raise KeyError(name)


def _build(name):
    N, qu, qd, ou, od, U, uv = _model(name)
    up, dn = H_TBq(N, qu), H_TBq(N, qd)
    up.add_ops(ou)
    dn.add_ops(od)
    return Hubbard(up, dn, U, Uvec=uv), (N, qu, qd, ou, od, U, uv)


@pytest.mark.parametrize("name", ["chain8_4_3", "kagome12_6_6", "kagome12_randomhop_uvec", "chain14_7_7",
                                  "chain8_0_3", "chain8_4_8", "chain8_8_0", "chain8_3_3_dn_nocalls"])
def test_slab_hv_bitwise(name):
    print(f"slab H.v {name}", flush=True)
    H, _ = _build(name)
    n = H.slab_layout(True)
    assert n >= H.dim
    x = torch.from_numpy(np.random.default_rng(3).normal(size=H.dim)).to(DEV)
    xl = H.to_slab(x)
    assert torch.equal(H.from_slab(xl), x)  # exact permutation
    yl = torch.full_like(xl, float("nan"))
    H.apply_slab(xl, yl)
    y = H.from_slab(yl)
    yref = H.apply(x)
    assert torch.equal(y, yref), float((y - yref).abs().max())
    # padding entries are zero and the layout holds every amplitude once
    assert torch.isfinite(yl).all()
    assert abs(float(yl.abs().sum()) - float(y.abs().sum())) <= 1e-9 * float(y.abs().sum())


def test_slab_lanczos_matches_reference_layout():
    print("slab Lanczos kagome12 6/6 U=4", flush=True)
    H, (N, qu, qd, ou, od, U, uv) = _build("kagome12_6_6")
    f = torch.empty(H.dim, dtype=torch.float64, device=DEV)
    ev = torch.empty_like(f)
    lz = Lanczos(H.dimUP, H.dimDN)
    lz.random_vector(f)
    e_ref = lz.eigen(H, f, ev)
    n_ref = lz.nsteps
    v_ref = ev.clone()
    H.slab_layout(True)
    lz2 = Lanczos(H.dimUP, H.dimDN)
    lz2.random_vector(f)
    e_slab = lz2.eigen(H, f, ev)
    assert lz2.nsteps == n_ref
    assert abs(e_slab - e_ref) <= 1e-10 * abs(e_ref)
    ov = abs(float(torch.dot(ev, v_ref)))
    assert abs(ov - 1.0) <= 1e-8
    # the rebuild (Ritz update fused into pass B) gives a unit eigenvector of H
    assert abs(float(torch.linalg.norm(ev)) - 1.0) <= 1e-12
    res = float(torch.linalg.norm(H.apply(ev) - e_slab * ev))
    assert res <= 1e-6 * abs(e_slab), res
    st, e0o, *_ = O.Hubbard(O.Species(N, qu, ou), O.Species(N, qd, od), U).lanczos_eigen(O.random_vector(H.dim))
    assert abs(e_slab - e0o) <= 1e-10 * abs(e0o)
    H.slab_layout(False)


def test_slab_tridiagonal_matches_reference_layout():
    """Lanczos<T>::tridiagonal on the slab layout: the same coefficients as on the reference
    layout up to the rounding of the two summation orders, and the oracle's to the bar of the
    reference-layout test (1e-9 absolute over the first 40 steps)"""
    print("slab tridiagonal kagome12 6/6 U=4, 200 steps", flush=True)
    H, (N, qu, qd, ou, od, U, uv) = _build("kagome12_6_6")
    f0 = O.random_vector(H.dim)
    f = torch.from_numpy(f0).to(DEV)
    l_ref = Lanczos(H.dimUP, H.dimDN)
    l_ref.tridiagonal(H, f, 200)
    H.slab_layout(True)
    l_slab = Lanczos(H.dimUP, H.dimDN)
    l_slab.tridiagonal(H, f, 200)
    assert torch.equal(f, torch.from_numpy(f0).to(DEV))  # d_f is not modified
    H.slab_layout(False)
    a, b = np.array(l_slab.a), np.array(l_slab.b)
    assert len(a) == 200 and len(b) == 201
    assert np.abs(a[:40] - np.array(l_ref.a[:40])).max() <= 1e-12
    assert np.abs(b[:41] - np.array(l_ref.b[:41])).max() <= 1e-12
    ao, bo = O.Hubbard(O.Species(N, qu, ou), O.Species(N, qd, od), U).lanczos_tridiagonal(f0, 200)
    assert np.allclose(a[:40], ao[:40], rtol=0, atol=1e-9)
    assert np.allclose(b[:41], bo[:41], rtol=0, atol=1e-9)
    # the lowest Ritz value of T_200 is converged: it is E0 on both layouts
    from scipy.linalg import eigvalsh_tridiagonal
    t_s = eigvalsh_tridiagonal(a, b[1:200], select="i", select_range=(0, 0))[0]
    t_r = eigvalsh_tridiagonal(np.array(l_ref.a), np.array(l_ref.b[1:200]), select="i", select_range=(0, 0))[0]
    assert abs(t_s - t_r) <= 1e-10 * abs(t_r)


def test_slab_rejects_complex_operator():
    up = H_TBq(6, 3, complex_tables=True)
    up.add_ops([(L.OP_HOP_PHASE, 0, 1, -1.0, 0.2)])
    H = Hubbard(up, up, 1.0)
    with pytest.raises(EdError):
        H.slab_layout(True)


GOLD = os.path.join(os.path.dirname(__file__), "golden", "kagome18y_lanczos100_oracle.json")


@pytest.mark.skipif(not os.path.exists(GOLD), reason="fixture not generated")
def test_slab_kagome18y_first_100_steps():
    """the 18-site fixture, with the slab layout built automatically by the solver"""
    gold = json.load(open(GOLD))
    n = gold["nsteps"]
    print("slab kagome18y 9/9 U=8: 100 Lanczos steps at D = 2.36e9", flush=True)
    hq = H_TBq(18, 9)
    hq.add_ops(L.hubbard_ops(L.KAGOME18Y, t=-1.0))
    H = Hubbard(hq, hq, 8.0)
    assert H.slab_len_query() == 0
    f = torch.empty(H.dim, dtype=torch.float64, device=DEV)
    lz = Lanczos(H.dimUP, H.dimDN, max_steps=n, check_every=1000)
    lz.random_vector(f)
    try:
        lz.eigen(H, f, None)
    except EdError:
        pass  # 100 steps do not converge: a and b are still returned
    # a real sector of 2^27 amplitudes and more gets the slab layout by itself
    assert H.slab_len_query() >= H.dim
    H.slab_layout(False)
    assert H.slab_len_query() == 0
    del f
    torch.cuda.empty_cache()
    a, b = np.array(lz.a), np.array(lz.b)
    ga, gb = np.array(gold["alpha"]), np.array(gold["beta"])
    print(f"max |d alpha| {np.abs(a - ga).max():.2e}, max |d beta| {np.abs(b[1:n + 1] - gb[1:n + 1]).max():.2e}",
          flush=True)
    assert len(a) == n
    assert np.abs(a - ga).max() < 1e-9 and np.abs(b[1:n + 1] - gb[1:n + 1]).max() < 1e-9
    from scipy.linalg import eigvalsh_tridiagonal
    theta = eigvalsh_tridiagonal(a, b[1:n], select="i", select_range=(0, 0))[0]
    assert abs(theta - gold["ritz_lowest"]) < 1e-10 * abs(gold["ritz_lowest"]), (theta, gold["ritz_lowest"])


@pytest.mark.parametrize("layout", ["reference", "slab"])
def test_lanczos_local_reorthogonalisation(layout):
    """the DGKS pass of f against v_N (ed_lanczos_cfg.reorth) on both layouts: E0 at the 1e-10 bar
    against the oracle, a unit Ritz vector with a small residual, and the recurrence it changes is
    still a Lanczos recurrence (the same E0 as without the pass, to 1e-10)"""
    print(f"Lanczos with reorth, kagome12 6/6 U=4, {layout} layout", flush=True)
    H, (N, qu, qd, ou, od, U, uv) = _build("kagome12_6_6")
    if layout == "slab":
        H.slab_layout(True)
    f = torch.empty(H.dim, dtype=torch.float64, device=DEV)
    ev = torch.empty_like(f)
    lz = Lanczos(H.dimUP, H.dimDN, reorth=True)
    lz.random_vector(f)
    e_re = lz.eigen(H, f, ev)
    lz0 = Lanczos(H.dimUP, H.dimDN)
    lz0.random_vector(f)
    e_plain = lz0.eigen(H, f, None)
    if layout == "slab":
        H.slab_layout(False)
    st, e0o, *_ = O.Hubbard(O.Species(N, qu, ou), O.Species(N, qd, od), U).lanczos_eigen(O.random_vector(H.dim))
    assert abs(e_re - e0o) <= 1e-10 * abs(e0o)
    assert abs(e_re - e_plain) <= 1e-10 * abs(e0o)
    assert abs(float(torch.linalg.norm(ev)) - 1.0) <= 1e-12
    res = float(torch.linalg.norm(H.apply(ev) - e_re * ev))
    assert res <= 1e-6 * abs(e_re), res


def test_slab_shapes_bitwise_and_autotune():
    """Every slab launch shape (tile rows x gathers in flight x tile order) gives the reference
    layout's H.v bit for bit, and the autotuner's choice is one of them."""
    print("slab shapes chain14", flush=True)
    H, _ = _build("chain14_7_7")
    H.slab_layout(True)
    x = torch.from_numpy(np.random.default_rng(5).normal(size=H.dim)).to(DEV)
    yref = H.apply(x)
    xl = H.to_slab(x)
    yl = torch.empty_like(xl)
    for tr in (32, 64):
        for nb in (8, 16):
            for order in (0, 2, 4, 8, 16):
                H.slab_set_shape(tr, nb, order)
                yl.fill_(float("nan"))
                H.apply_slab(xl, yl)
                assert torch.equal(H.from_slab(yl), yref), (tr, nb, order)
    # the fused Lanczos dot (one partial per 32-row unit) is the same for every shape: the
    # recurrence coefficients agree bit for bit
    coef = []
    f0 = torch.from_numpy(O.random_vector(H.dim)).to(DEV)
    for tr, nb, order in ((32, 8, 8), (64, 16, 0), (64, 8, 4)):
        H.slab_set_shape(tr, nb, order)
        lz = Lanczos(H.dimUP, H.dimDN)
        lz.tridiagonal(H, f0, 20)
        coef.append(np.array(lz.a + lz.b))
    for c in coef[1:]:
        assert np.array_equal(c.view(np.uint64), coef[0].view(np.uint64))
    best = H.slab_autotune(xl, yl, reps=1)
    assert best["tile_rows"] in (32, 64) and best["gathers_in_flight"] in (8, 16)
    assert best["tile_order"] in (0, 2, 4, 8, 16) and best["ms_per_apply"] > 0
    assert abs(best["ms_per_apply"] - min(v for v in best["candidate_ms"] if v > 0)) <= 1e-3
    H.apply_slab(xl, yl)
    assert torch.equal(H.from_slab(yl), yref)
    with pytest.raises(EdError):
        H.slab_set_shape(48, 8, 8)


def test_slab_buffer_checks():
    """The layout conversions and H.v check their buffers against the handle's layout length and
    dtype before any kernel runs."""
    H, _ = _build("kagome12_6_6")
    n = H.slab_layout(True)
    x = torch.zeros(H.dim, dtype=torch.float64, device=DEV)
    with pytest.raises(ValueError):
        H.to_slab(x, torch.empty(n - 1, dtype=torch.float64, device=DEV))
    with pytest.raises(ValueError):
        H.to_slab(x.to(torch.complex128))
    with pytest.raises(ValueError):
        H.from_slab(torch.empty(n + 16, dtype=torch.float64, device=DEV))
    with pytest.raises(ValueError):
        H.from_slab(torch.empty(n, dtype=torch.float32, device=DEV))
    xl = H.to_slab(x)
    with pytest.raises(ValueError):
        H.apply_slab(xl, xl)
    H.slab_layout(False)
    with pytest.raises(ValueError):
        H.to_slab(x)

"""GPU parity of the core path against the CPU oracle: basis, tables, start vector,
H.v (real, complex, per-site U, dilution mask), BLAS-1, Lanczos, tridiagonalisation,
S^2 -- all through the C ABI of libedhip.so (ctypes), with torch only holding device
buffers.

Bars: integer tables and the start vector bit-exact; H.v within 1e-14 of max|y|
(same per-element summation order, FMA contraction differs); E0 within 1e-10 relative
(the parity bar), same number of Lanczos steps, Ritz vectors parallel to 1e-8.
"""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from exact_diagonalization_for_fermionic_systems_with_gpu_amd import lattices as L
from exact_diagonalization_for_fermionic_systems_with_gpu_amd import (H_TBq, Hubbard, Lanczos, Thrust_Lin,
                                                                       compute_sigma_omega)
from exact_diagonalization_for_fermionic_systems_with_gpu_amd.totalspin import TotalSpinHubbard

pytestmark = pytest.mark.gpu

DEV = "cuda"


def gpu_species(N, Q, ops, cplx=False):
    h = H_TBq(N, Q, complex_tables=cplx)
    h.add_ops(ops)
    return h


def to_np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("N,Q", [(12, 6), (14, 7), (18, 9), (12, 0), (12, 12), (4, None), (20, 3)])
def test_basis_bit_exact(N, Q):
    print(f"basis N={N} Q={Q}", flush=True)
    g = H_TBq(N, Q)
    o = O.Species(N, -1 if Q is None else Q)
    assert g.count == o.count
    assert np.array_equal(to_np(g.state), o.state)
    assert np.array_equal(to_np(g.address), o.address)


def _table_cases():
    r = np.random.default_rng(7).uniform(0.8, 1.2, 12)
    kx, ky = L.kinetic_ops(L.KAGOME12, r)
    jx, jy = L.current_ops(L.KAGOME12, r)
    return {
        "kagome12_randomhop": (12, 6, L.hubbard_ops(L.KAGOME12, r=r)),
        "kagome12_kx": (12, 6, kx), "kagome12_jy": (12, 5, jy), "kagome12_jx": (12, 6, jx),
        "kagome18y": (18, 9, L.hubbard_ops(L.KAGOME18Y, t=-1.0)),
        "lieb_dup": (3, 1, [(0, 0, 1, -1.0), (0, 1, 0, -1.0), (0, 0, 2, -1.0), (0, 2, 0, -1.0)]),
        "zero_slot": (6, 3, [(0, 0, 1, 0.0), (0, 1, 2, -1.0), (2, 3, 3, 0.5), (3, 1, 4, 0.25), (0, 4, 5, 1.0)]),
        "full_fock": (4, None, [(0, 0, 1, -1.0), (0, 2, 3, -0.5), (2, 1, 1, 2.0)]),
    }


@pytest.mark.parametrize("name", list(_table_cases().keys()))
def test_tables_bit_exact(name):
    N, Q, ops = _table_cases()[name]
    print(f"tables {name}", flush=True)
    g = gpu_species(N, Q, ops)
    o = O.Species(N, -1 if Q is None else Q, ops)
    assert np.array_equal(to_np(g.Ax).view(np.uint64), o.Ax.view(np.uint64))
    assert np.array_equal(to_np(g.Aj), o.Aj)


def test_complex_phase_tables():
    ops = L.lieb_twisted_ops(2 * math.pi * 3 / 16)
    g = gpu_species(12, 5, ops, cplx=True)
    o = O.Species(12, 5, ops, complex_tables=True)
    assert np.array_equal(to_np(g.Aj), o.Aj)
    assert np.allclose(to_np(g.Ax), o.Ax, rtol=0, atol=4e-16)


def test_slot_overflow_raises():
    g = gpu_species(3, 1, [(2, 0, 0, 1.0 + k) for k in range(12)])
    with pytest.raises(RuntimeError):
        g.Ax


@pytest.mark.parametrize("cplx", [False, True])
def test_random_vector_bit_exact(cplx):
    n = 853776
    t = torch.empty(n, dtype=torch.complex128 if cplx else torch.float64, device=DEV)
    from exact_diagonalization_for_fermionic_systems_with_gpu_amd._lib import call, dptr, stream_ptr
    call("ed_random_vector", n, 235, int(cplx), dptr(t), 0, stream_ptr())
    o = O.random_vector(n, complex_=cplx, normalize=False)
    assert np.array_equal(to_np(t).view(np.uint64), o.view(np.uint64))
    lz = Lanczos(1, n)
    lz.random_vector(t)
    on = O.random_vector(n, complex_=cplx, normalize=True)
    assert np.max(np.abs(to_np(t) - on)) < 1e-17


def _hv_cases():
    r = np.random.default_rng(11).uniform(0.8, 1.2, 12)
    eps = np.random.default_rng(5).uniform(-8, 8, 14)
    chain = L.hubbard_ops(L.chain(14), t=-1.0) + [(2, i, i, float(v)) for i, v in enumerate(eps)]
    return {
        "kagome12_6_6_U4": (12, 6, 6, L.hubbard_ops(L.KAGOME12, t=-1.0), 4.0, None, None, False),
        "kagome12_randomhop_5_7": (12, 5, 7, L.hubbard_ops(L.KAGOME12, r=r), 8.0, None, None, False),
        "chain14_disorder": (14, 7, 7, chain, 4.0, None, None, False),
        "lieb12_diluted_7_5": (12, 7, 5, L.hubbard_ops(L.LIEB12, t=-1.0), 4.0, 1755, None, False),
        "lieb12_uvec_6_6": (12, 6, 6, L.hubbard_ops(L.LIEB12, t=-1.0), 0.0, None,
                            np.random.default_rng(2).uniform(2, 6, 12), False),
        "lieb12_twist_4_8": (12, 4, 8, L.lieb_twisted_ops(2 * math.pi * 5 / 16), 4.0, None, None, True),
        "dim1_sector": (12, 12, 0, L.hubbard_ops(L.LIEB12, t=-1.0), 4.0, None, None, False),
        "narrow_dn_3_1": (12, 3, 1, L.hubbard_ops(L.LIEB12, t=-1.0), 2.0, None, None, False),
    }


_OTF_ENV = {"otf": {}, "otf_w8": {"ED_HV_W": "8"}, "otf_w16": {"ED_HV_W": "16"}, "otf_w32": {"ED_HV_W": "32"},
            "otf_w64": {"ED_HV_W": "64"}, "otf_w256": {"ED_HV_W": "256"}, "otf_w512": {"ED_HV_W": "512"},
            "otf_rcm": {"ED_HV_UPORDER": "1"}, "otf_rcm_w16": {"ED_HV_UPORDER": "1", "ED_HV_W": "16"},
            "otf_rcm_w32": {"ED_HV_UPORDER": "1", "ED_HV_W": "32"},
            "otf_notab": {"ED_HV_UPTAB": "0"}, "otf_notab_rcm_w16": {"ED_HV_UPTAB": "0", "ED_HV_UPORDER": "1",
                                                                  "ED_HV_W": "16"},
            "otf_nomask_r4": {"ED_HV_DNMASK": "0", "ED_HV_RMAX": "4"}, "otf_r1": {"ED_HV_RMAX": "1"},
            "otf_dn1": {"ED_HV_DN2": "0"}, "otf_k4": {"ED_HV_DNK": "4"}, "otf_k1": {"ED_HV_DNK": "1"},
            "otf_k1_r4": {"ED_HV_DNK": "1", "ED_HV_RMAX": "4"}}


@pytest.mark.parametrize("path", list(_OTF_ENV) + ["tables", "rowkernel"])
@pytest.mark.parametrize("name", list(_hv_cases().keys()))
def test_hv_parity(name, path, monkeypatch):
    """otf*: descriptors generated from the call list (default W = 128 columns per UP slab, DN
    descriptors cached over 2 panels of 8 rows; the variants change the slab width -- 16 B loads
    per lane from W = 16 when dim_dn is even -- the UP row order (reverse Cuthill-McKee), the DN
    panels and the masking); tables: two-pass kernels on the ELL tables; rowkernel: the
    single-pass row kernel (ED_HV_ALGO=1)."""
    N, Qu, Qd, ops, U, mask, Uvec, cplx = _hv_cases()[name]
    print(f"hv {name} {path}", flush=True)
    gu, gd = gpu_species(N, Qu, ops, cplx), gpu_species(N, Qd, ops, cplx)
    if path == "rowkernel":
        monkeypatch.setenv("ED_HV_ALGO", "1")
    for k, v in _OTF_ENV.get(path, {}).items():
        monkeypatch.setenv(k, v)
    H = Hubbard(gu, gd, U, dilution=mask, Uvec=Uvec, from_tables=not path.startswith("otf"))
    ou, od = O.Species(N, Qu, ops, cplx), O.Species(N, Qd, ops, cplx)
    OH = O.Hubbard(ou, od, U, mask=mask, Uvec=Uvec)
    rng = np.random.default_rng(1)
    x = rng.normal(size=OH.dim)
    if cplx:
        x = x + 1j * rng.normal(size=OH.dim)
    xg = torch.from_numpy(x).to(DEV)
    y = to_np(H.apply(xg))
    yo = OH.apply(x)
    scale = max(np.max(np.abs(yo)), 1e-300)
    assert np.max(np.abs(y - yo)) <= 1e-14 * scale
    if not cplx:
        # real vectors through the complex path (Hubbard_mv<complex<double>>) give the same numbers
        yc = to_np(H.apply(xg.to(torch.complex128)))
        assert np.max(np.abs(yc.real - y)) <= 1e-14 * scale and np.max(np.abs(yc.imag)) == 0.0


def test_hv_hermitian_and_linear():
    gu = gpu_species(12, 6, L.hubbard_ops(L.KAGOME12, t=-1.0))
    H = Hubbard(gu, gu, 4.0)
    g = torch.Generator(device="cpu").manual_seed(3)
    x = torch.randn(H.dim, generator=g, dtype=torch.float64).to(DEV)
    z = torch.randn(H.dim, generator=g, dtype=torch.float64).to(DEV)
    Hx, Hz = H.apply(x), H.apply(z)
    assert abs(float(torch.dot(z, Hx) - torch.dot(Hz, x))) < 1e-10 * float(torch.linalg.norm(Hx))
    Hs = H.apply(2.5 * x - z)
    assert float(torch.max(torch.abs(Hs - (2.5 * Hx - Hz)))) < 1e-12 * float(torch.max(torch.abs(Hs)))


def test_blas_against_torch():
    n = 100003
    g = torch.Generator(device="cpu").manual_seed(4)
    a = torch.randn(n, generator=g, dtype=torch.float64).to(DEV)
    b = torch.randn(n, generator=g, dtype=torch.float64).to(DEV)
    tl = Thrust_Lin(1, n)
    assert abs(tl.thrust_Dotprod(a, b) - float(torch.dot(a, b))) < 1e-11
    assert abs(tl.norm(a) - float(torch.dot(a, a))) < 1e-9
    c = tl.axpy(-0.5, a, b)
    assert torch.allclose(c, b - 0.5 * a, rtol=0, atol=1e-15)
    ac = (a + 1j * b).contiguous()
    bc = (b - 0.3j * a).contiguous()
    d = tl.thrust_Dotprod(ac, bc)
    ref = complex(torch.vdot(ac, bc).cpu())
    assert abs(d - ref) < 1e-9


def test_lanczos_parity_kagome12():
    print("lanczos kagome12 6/6 U=4", flush=True)
    ops = L.hubbard_ops(L.KAGOME12, t=-1.0)
    gu = gpu_species(12, 6, ops)
    H = Hubbard(gu, gu, 4.0)
    ou = O.Species(12, 6, ops)
    OH = O.Hubbard(ou, ou, 4.0)
    st, e0o, evo, ao, bo, no = OH.lanczos_eigen(O.random_vector(OH.dim))
    lanc = Lanczos(gu.count, gu.count)
    f = torch.empty(H.dim, dtype=torch.float64, device=DEV)
    ev = torch.empty_like(f)
    lanc.random_vector(f)
    e0 = lanc.eigen(H, f, ev)
    assert lanc.nsteps == no
    assert abs(e0 - e0o) <= 1e-10 * abs(e0o)
    # the recurrence has no reorthogonalisation, so rounding differences (FMA contraction,
    # reduction order) grow once Ritz values converge; the early coefficients agree tightly
    assert np.allclose(lanc.a[:30], ao[:30], rtol=0, atol=1e-10)
    evg = to_np(ev)
    assert abs(abs(np.dot(evg, evo)) - 1.0) < 1e-8
    r = to_np(H.apply(ev)) - e0 * evg
    assert np.linalg.norm(r) < 1e-4


def test_lanczos_complex_twist_and_breakdown():
    ops = L.lieb_twisted_ops(2 * math.pi * 3 / 16)
    gu, gd = gpu_species(12, 5, ops, True), gpu_species(12, 7, ops, True)
    H = Hubbard(gu, gd, 4.0)
    OH = O.Hubbard(O.Species(12, 5, ops, True), O.Species(12, 7, ops, True), 4.0)
    st, e0o, *_ = OH.lanczos_eigen(O.random_vector(OH.dim, complex_=True))
    f = torch.empty(H.dim, dtype=torch.complex128, device=DEV)
    ev = torch.empty_like(f)
    lanc = Lanczos(gu.count, gd.count)
    lanc.random_vector(f)
    e0 = lanc.eigen(H, f, ev)
    assert abs(e0 - e0o) <= 1e-10 * abs(e0o)
    # dimension-1 sector: exact breakdown at step 0
    g1, g0 = gpu_species(12, 12, ops, True), gpu_species(12, 0, ops, True)
    H1 = Hubbard(g1, g0, 4.0)
    f1 = torch.empty(1, dtype=torch.complex128, device=DEV)
    e1 = torch.empty_like(f1)
    l1 = Lanczos(1, 1)
    l1.random_vector(f1)
    assert abs(l1.eigen(H1, f1, e1)) < 1e-14 and l1.nsteps == 1


def test_tridiagonal_and_continued_fraction():
    ops = L.hubbard_ops(L.KAGOME12, t=-1.0)
    gu = gpu_species(12, 6, ops)
    H = Hubbard(gu, gu, 4.0)
    OH = O.Hubbard(O.Species(12, 6, ops), O.Species(12, 6, ops), 4.0)
    f0 = O.random_vector(OH.dim)
    ao, bo = OH.lanczos_tridiagonal(f0, 200)
    lanc = Lanczos(gu.count, gu.count)
    lanc.tridiagonal(H, torch.from_numpy(f0).to(DEV), 200)
    # exact arithmetic would keep these equal; rounding in the unreorthogonalised recurrence
    # grows with the step count, so compare the first steps tightly, all loosely
    assert np.allclose(lanc.a[:40], ao[:40], rtol=0, atol=1e-9)
    assert np.allclose(lanc.b[:41], bo[:41], rtol=0, atol=1e-9)
    w = 0.004 * (np.arange(8000) + 1.0) + 0.01j
    from exact_diagonalization_for_fermionic_systems_with_gpu_amd.hubmatrix import correlator_continued_fraction
    import ctypes as C
    from exact_diagonalization_for_fermionic_systems_with_gpu_amd._lib import call
    g = np.empty_like(w)
    a = np.ascontiguousarray(ao)
    b = np.ascontiguousarray(bo)
    call("ed_continued_fraction", len(a), a.ctypes.data_as(C.POINTER(C.c_double)),
         b.ctypes.data_as(C.POINTER(C.c_double)), -10.5, 2.0, len(w), w.ctypes.data_as(C.POINTER(C.c_double)),
         g.ctypes.data_as(C.POINTER(C.c_double)))
    go = O.continued_fraction(a, b, -10.5, 2.0, w)
    assert np.max(np.abs(g - go)) <= 1e-12 * np.max(np.abs(go))


def test_s2_parity_and_singlet():
    ops = L.hubbard_ops(L.KAGOME12, t=-1.0)
    gu, gd = gpu_species(12, 5, ops), gpu_species(12, 6, ops)
    S = TotalSpinHubbard(gu, gd)
    x = np.random.default_rng(9).normal(size=gu.count * gd.count)
    y = to_np(S.Hubbard_mv_new(torch.from_numpy(x).to(DEV)))
    yo = O.s2_apply(O.Species(12, 5, ops), O.Species(12, 6, ops), x)
    assert np.max(np.abs(y - yo)) <= 1e-14 * np.max(np.abs(yo))
    # ground states by the GPU Lanczos, total spin by the GPU S^2 kernel: the 12-electron
    # ground state (U = 4) is a singlet, the 11-electron one a doublet, S(S+1) = S_z^2 + S_xy
    for qu, ss_expect in ((6, 0.0), (5, 0.75)):
        hu = gpu_species(12, qu, ops)
        H = Hubbard(hu, gd, 4.0)
        f = torch.empty(H.dim, dtype=torch.float64, device=DEV)
        ev = torch.empty_like(f)
        lz = Lanczos(hu.count, gd.count)
        lz.random_vector(f)
        lz.eigen(H, f, ev)
        ev /= torch.linalg.vector_norm(ev)
        S2 = TotalSpinHubbard(hu, gd)
        sz = 0.5 * (qu - 6)
        ss = sz * sz + float(torch.dot(ev, S2.Hubbard_mv_new(ev)).item())
        assert abs(ss - ss_expect) <= 1e-8, (qu, ss)


@pytest.mark.parametrize("order", ["0", "1"])
def test_autotune_keeps_results_bitwise(order, monkeypatch):
    """The tile-shape autotuner picks among kernels that add the same terms in the same order:
    H.v after tuning equals H.v before, bit for bit, for real and complex vectors (order 1: the
    UP row-order candidates are in the search)."""
    monkeypatch.setenv("ED_HV_UPORDER", order)
    ops = L.hubbard_ops(L.KAGOME12, t=-1.0)
    gu, gd = gpu_species(12, 6, ops), gpu_species(12, 6, ops)
    H = Hubbard(gu, gd, 4.0)
    x = torch.from_numpy(np.random.default_rng(3).normal(size=H.dim)).to(DEV)
    y0 = H.apply(x)
    tmp = torch.empty_like(x)
    best = H.autotune(x, tmp, reps=2)
    assert best["slab_width"] in (32, 64, 128, 256) and best["ms_per_apply"] > 0
    y1 = H.apply(x)
    assert torch.equal(y0, y1)
    xc = x.to(torch.complex128) * (1 + 0.5j)
    assert torch.equal(H.apply(xc), torch.complex(y0, 0.5 * y0))

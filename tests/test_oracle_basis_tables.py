"""Pins of the oracle's basis, sign and ELL-table restatement (fermi_operator.h:19-149).

* bit-exact against the reference's own fermi_operator.h compiled into oracle/_ref
  (skipped where the reference tree is absent);
* closed forms: colex rank = sum_k C(p_k, k+1), count = C(N,Q), the JW sign as a popcount
  of the between-mask, exhaustively;
* an independent second-quantised (Jordan-Wigner) dense construction of the hopping
  operator, which a dropped sign or transposed index in the table builder would fail.
"""
import ctypes as C
import itertools
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from exact_diagonalization_for_fermionic_systems_with_gpu_amd import lattices as L

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SO = os.path.join(REPO, "oracle", "_ref", "libref_fermi.so")


def _ref():
    if not os.path.exists(REF_SO):
        if os.path.isdir("/root/reference"):
            os.system(f"make -s -C {os.path.join(REPO, 'oracle')} ref >/dev/null 2>&1")
    if not os.path.exists(REF_SO):
        pytest.skip("reference table builder (oracle/_ref) not built")
    R = C.CDLL(REF_SO)
    R.ref_htbq_tables.restype = C.c_longlong
    R.ref_htbq_tables.argtypes = [C.c_int] * 3 + [C.c_void_p] * 9
    R.ref_count1_bt_sites.restype = C.c_int
    R.ref_count1_bt_sites.argtypes = [C.c_int] * 3
    return R


def ref_tables(N, Q, ops):
    R = _ref()
    k = np.array([o[0] for o in ops], np.int32)
    i = np.array([o[1] for o in ops], np.int32)
    j = np.array([o[2] for o in ops], np.int32)
    v = np.array([o[3] for o in ops], np.float64)
    dg = np.array([o[5] if len(o) > 5 else 0 for o in ops], np.int32)
    cnt = math.comb(N, Q) if Q >= 0 else (1 << N)
    S = 10 * (Q if Q >= 0 else N) + 1
    st = np.zeros(cnt, np.int32)
    ad = np.zeros(1 << N, np.int32)
    Ax = np.zeros(cnt * S)
    Aj = np.zeros(cnt * S, np.int32)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    n = R.ref_htbq_tables(N, Q, len(ops), p(k), p(i), p(j), p(v), p(st), p(ad), p(Ax), p(Aj), p(dg))
    assert n == cnt
    return st, ad, Ax, Aj


def test_count1_bt_exhaustive_closed_form():
    for n in range(1 << 10):
        for i in range(10):
            for j in range(10):
                if i == j:
                    continue
                lo, hi = min(i, j), max(i, j)
                mask = ((1 << (hi - lo - 1)) - 1) << (lo + 1)
                assert O.count1_bt(n, i, j) == bin(n & mask).count("1")


def test_count1_bt_matches_reference_code():
    R = _ref()
    for n in range(0, 1 << 12, 7):
        for i in range(12):
            for j in range(12):
                if i != j:
                    assert R.ref_count1_bt_sites(n, i, j) == O.count1_bt(n, i, j)


@pytest.mark.parametrize("N,Q", [(4, 2), (12, 6), (14, 7), (10, 0), (10, 10), (18, 9)])
def test_basis_closed_form(N, Q):
    sp = O.Species(N, Q)
    assert sp.count == math.comb(N, Q)
    st = sp.state.astype(np.int64)
    assert np.all(np.diff(st) > 0)
    pc = np.array([bin(int(s)).count("1") for s in st[:2000]])
    assert np.all(pc == Q)
    # colex rank closed form
    for r in list(range(min(50, sp.count))) + [sp.count - 1]:
        s = int(st[r])
        bits = [p for p in range(N) if (s >> p) & 1]
        assert sum(math.comb(p, k + 1) for k, p in enumerate(bits)) == r
        assert sp.address[s] == r
    # address is 0 outside the sector
    outside = [n for n in range(min(1 << N, 5000)) if bin(n).count("1") != Q]
    assert np.all(sp.address[outside] == 0)


def _ops_sets():
    k12 = L.KAGOME12
    r = np.random.default_rng(7).uniform(0.8, 1.2, 12)
    kx, ky = L.kinetic_ops(k12, r)
    jx, jy = L.current_ops(k12, r)
    lieb_pi = [(L.OP_HOP, i, j, (1.0 if k in L.LIEB12_PI_X_BONDS else -1.0)) for k, (i, j) in
               enumerate(L.LIEB12.bonds)]
    dup = [(L.OP_HOP, 0, 1, -1.0), (L.OP_HOP, 1, 0, -1.0), (L.OP_HOP, 0, 2, -1.0), (L.OP_HOP, 2, 0, -1.0)]
    zero = [(L.OP_HOP, 0, 1, 0.0), (L.OP_HOP, 1, 2, -1.0), (L.OP_DENSITY, 3, 3, 0.5),
            (L.OP_DENSITY2, 1, 4, 0.25), (L.OP_HOP, 4, 5, 1.0)]
    return {
        "kagome12_randomhop": (12, 6, L.hubbard_ops(k12, r=r)),
        "kagome12_kin_x": (12, 6, kx), "kagome12_kin_y": (12, 5, ky),
        "kagome12_cur_x": (12, 6, jx), "kagome12_cur_y": (12, 7, jy),
        "lieb12_pi": (12, 5, lieb_pi),
        "lieb3_duplicate_bonds": (3, 1, dup),
        "zero_valued_slot": (6, 3, zero),
        "chain14_disorder": (14, 7, L.hubbard_ops(L.chain(14), t=-1.0) +
                             [(L.OP_DENSITY, i, i, float(v)) for i, v in
                              enumerate(np.random.default_rng(3).uniform(-8, 8, 14))]),
        "full_fock_4": (4, -1, [(L.OP_HOP, 0, 1, -1.0), (L.OP_HOP, 2, 3, -0.5), (L.OP_DENSITY, 1, 1, 2.0)]),
        # Insert_1Fop / Insert_2Fop, every sign combination and both site orders, mixed with hops
        "full_fock_6_fermi_ops": (6, -1, [(O.OP_FERMI1, 2, 2, 0.7, 0.0, 0), (O.OP_FERMI1, 4, 4, -1.3, 0.0, 1),
                                          (O.OP_FERMI2, 1, 4, 0.5, 0.0, 2), (O.OP_FERMI2, 5, 0, 1.1, 0.0, 1),
                                          (O.OP_FERMI2, 3, 1, 0.9, 0.0, 0), (O.OP_FERMI2, 0, 5, -0.4, 0.0, 3),
                                          (L.OP_HOP, 0, 1, -1.0), (O.OP_FERMI1, 5, 5, 0.25, 0.0, 0),
                                          (O.OP_FERMI2, 2, 3, 0.3, 0.0, 1), (L.OP_DENSITY, 3, 3, 0.5)]),
        # number-changing calls on a sector basis ("use only full Hilbert space"): the target's
        # address is 0 outside the sector, and the reference stores that column as it is
        "sector_fermi_misuse": (6, 3, [(O.OP_FERMI1, 1, 1, 1.0, 0.0, 1), (O.OP_FERMI2, 0, 2, 0.5, 0.0, 1)]),
    }


@pytest.mark.parametrize("name", list(_ops_sets().keys()))
def test_tables_bit_exact_vs_reference_code(name):
    N, Q, ops = _ops_sets()[name]
    st, ad, Ax, Aj = ref_tables(N, Q, ops)
    sp = O.Species(N, Q, ops)
    assert np.array_equal(sp.state, st)
    assert np.array_equal(sp.address, ad)
    assert np.array_equal(sp.Ax.view(np.uint64), Ax.view(np.uint64))
    assert np.array_equal(sp.Aj, Aj)


def jw_dense_hop(N, Q, bonds_t):
    """Independent construction: c_k = (prod_{l<k} Z_l) a_k on the 2^N Fock space,
    H = sum t (c+_i c_j + c+_j c_i), restricted to popcount Q (ascending order)."""
    dim = 1 << N
    H = np.zeros((dim, dim))
    for (i, j, t) in bonds_t:
        for n in range(dim):
            for (a, b) in ((i, j), (j, i)):  # c+_a c_b
                if (n >> b) & 1 and not (n >> a) & 1:
                    s1 = bin(n & ((1 << b) - 1)).count("1")
                    m = n ^ (1 << b)
                    s2 = bin(m & ((1 << a) - 1)).count("1")
                    H[m | (1 << a), n] += t * (-1) ** (s1 + s2)
    idx = [n for n in range(dim) if bin(n).count("1") == Q]
    return H[np.ix_(idx, idx)]


@pytest.mark.parametrize("N,Q", [(6, 3), (8, 3), (7, 2)])
def test_tables_vs_jordan_wigner(N, Q):
    rng = np.random.default_rng(N * 10 + Q)
    bonds = []
    for _ in range(10):
        i, j = rng.choice(N, 2, replace=False)
        bonds.append((int(i), int(j), float(rng.normal())))
    sp = O.Species(N, Q, [(L.OP_HOP, i, j, t) for (i, j, t) in bonds])
    M = np.zeros((sp.count, sp.count))
    S = sp.slots
    for s in range(S):
        for a in range(sp.count):
            v = sp.Ax[s * sp.count + a]
            if v != 0:
                M[a, sp.Aj[s * sp.count + a]] += v
    ref = jw_dense_hop(N, Q, bonds)
    assert np.allclose(M, ref, atol=1e-14)
    assert np.allclose(M, M.T)


def test_single_particle_sector_is_hopping_matrix():
    cl = L.KAGOME12
    sp = O.Species(12, 1, L.hubbard_ops(cl, t=-1.0))
    M = np.zeros((12, 12))
    for s in range(sp.slots):
        for a in range(12):
            v = sp.Ax[s * 12 + a]
            if v:
                M[a, sp.Aj[s * 12 + a]] += v
    T = np.zeros((12, 12))
    for (i, j) in cl.bonds:
        T[i, j] += -1.0
        T[j, i] += -1.0
    # Q=1: state index a is site a (1 << a ascending); no JW signs
    assert np.array_equal(M, T)
    # kagome flat band at +2 (for t=-1 matrix elements) with multiplicity >= 3 on this cluster
    w = np.linalg.eigvalsh(T)
    assert np.isclose(w.max(), 2.0) and np.sum(np.isclose(w, 2.0)) >= 3


def test_slot_overflow_is_reported():
    # 12 density terms on one site exceed the 10*Q+1 = 11 slots of Q=1
    ops = [(L.OP_DENSITY, 0, 0, 1.0 + k) for k in range(12)]
    with pytest.raises(RuntimeError):
        O.Species(3, 1, ops)

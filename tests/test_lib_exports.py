"""CPU checks of the C ABI: libedhip.so loads and exports every function include/edhip.h
declares, and the ctypes binding declares the same set."""
import ctypes
import os
import re

from exact_diagonalization_for_fermionic_systems_with_gpu_amd import _lib

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(REPO, "include", "edhip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return set(re.findall(r"\b(ed_[a-z0-9_]+)\s*\(", txt))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = header_functions()
    assert len(names) >= 20
    for n in sorted(names):
        assert hasattr(L, n), f"{n} missing from libedhip.so"


def test_binding_matches_header():
    assert header_functions() == set(_lib.SIGNATURES)


def test_host_only_entry_points():
    import numpy as np
    L = _lib.lib()
    cnt = ctypes.c_int64()
    assert L.ed_basis_count(18, 9, ctypes.byref(cnt)) == 0 and cnt.value == 48620
    assert L.ed_basis_count(40, 3, ctypes.byref(cnt)) != 0
    assert b"bad" in L.ed_last_error()
    d = np.array([2.0, -1.0, 0.5, 3.0])
    e = np.array([1.0, 0.3, 0.7])
    lam = ctypes.c_double()
    z = np.zeros(4)
    P = ctypes.POINTER(ctypes.c_double)
    assert L.ed_tridiag_lowest(4, d.ctypes.data_as(P), e.ctypes.data_as(P), ctypes.byref(lam),
                               z.ctypes.data_as(P)) == 0
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    w, V = np.linalg.eigh(T)
    assert abs(lam.value - w[0]) < 1e-13
    assert np.allclose(np.abs(z), np.abs(V[:, 0]), atol=1e-12)
    ww = np.zeros(4)
    assert L.ed_tridiag_eigvals(4, d.ctypes.data_as(P), e.ctypes.data_as(P), ww.ctypes.data_as(P)) == 0
    assert np.allclose(ww, w, atol=1e-13)

"""ctypes front end of the CPU oracle (oracle/ed_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  Every function here
forwards to the C restatement of the reference arithmetic and returns numpy
arrays; the reference file:line each one follows is cited in ed_oracle.c.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

Z_SLOTS = 10  # fermi_operator.h:11

OP_HOP, OP_CURRENT, OP_DENSITY, OP_DENSITY2, OP_HOP_PHASE, OP_FERMI1, OP_FERMI2 = 0, 1, 2, 3, 4, 5, 6

_p = C.c_void_p
_i64 = C.c_int64


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.or_count1.restype = C.c_int
        L.or_count1.argtypes = [_i64]
        L.or_count1_bt.restype = C.c_int
        L.or_count1_bt.argtypes = [_i64, C.c_int, C.c_int]
        L.or_basis.restype = _i64
        L.or_basis.argtypes = [C.c_int, C.c_int, _p, _p]
        L.or_tables.restype = C.c_int
        L.or_tables.argtypes = [C.c_int, C.c_int, _p, _p, _i64, C.c_int, _p, _p, _p, _p, _p, _p, _p]
        L.or_tables_z.restype = C.c_int
        L.or_tables_z.argtypes = [C.c_int, C.c_int, _p, _p, _i64, C.c_int, _p, _p, _p, _p, _p, _p, _p, _p]
        L.or_hubbard_apply.restype = None
        L.or_hubbard_apply.argtypes = [_i64, _i64, C.c_int, _p, _p, C.c_int, _p, _p, _p, _p, C.c_double, _i64,
                                       _p, _p, _p]
        L.or_hubbard_apply_rows.restype = None
        L.or_hubbard_apply_rows.argtypes = L.or_hubbard_apply.argtypes + [_i64, _i64]
        L.or_hubbard_apply_z.restype = None
        L.or_hubbard_apply_z.argtypes = L.or_hubbard_apply.argtypes
        L.or_s2_apply.restype = None
        L.or_s2_apply.argtypes = [C.c_int, _i64, _i64, _p, _p, _p, _p, _p, _p]
        L.or_random_vector.restype = None
        L.or_random_vector.argtypes = [_i64, C.c_uint32, _p]
        L.or_random_vector_z.restype = None
        L.or_random_vector_z.argtypes = [_i64, C.c_uint32, _p]
        L.or_dot.restype = C.c_double
        L.or_dot.argtypes = [_i64, _p, _p]
        L.or_normalize.restype = None
        L.or_normalize.argtypes = [_i64, _p]
        L.or_normalize_z.restype = None
        L.or_normalize_z.argtypes = [_i64, _p]
        L.or_tridiag_lowest.restype = C.c_double
        L.or_tridiag_lowest.argtypes = [C.c_int, _p, _p, _p]
        L.or_tridiag_eigvals.restype = None
        L.or_tridiag_eigvals.argtypes = [C.c_int, _p, _p, _p]
        L.or_op_new.restype = _p
        L.or_op_new.argtypes = [_i64, _i64, C.c_int, _p, _p, C.c_int, _p, _p, _p, _p, C.c_double, _i64, _p, C.c_int]
        L.or_op_free.restype = None
        L.or_op_free.argtypes = [_p]
        L.or_op_apply.restype = None
        L.or_op_apply.argtypes = [_p, _p, _p]
        L.or_lanczos_eigen.restype = C.c_int
        L.or_lanczos_eigen.argtypes = [_p, C.c_int, C.c_double, _p, _p, _p, _p, _p, _p]
        L.or_lanczos_lowest_2vec.restype = C.c_int
        L.or_lanczos_lowest_2vec.argtypes = [_p, C.c_int, C.c_double, C.c_uint32, _p, _p, _p, _p, C.c_int]
        L.or_lanczos_tridiagonal.restype = C.c_int
        L.or_lanczos_tridiagonal.argtypes = [_p, C.c_int, _p, _p, _p]
        L.or_continued_fraction.restype = None
        L.or_continued_fraction.argtypes = [C.c_int, _p, _p, C.c_double, C.c_double, C.c_int, _p, _p]
        L.or_drude_combine.restype = C.c_double
        L.or_drude_combine.argtypes = [C.c_int, _p, _p, C.c_double, C.c_double, C.c_double, C.c_double]
        L.or_num_threads.restype = C.c_int
        _LIB = L
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def count1(s):
    return lib().or_count1(int(s))


def count1_bt(n, i, j):
    return lib().or_count1_bt(int(n), int(i), int(j))


class Species:
    """One spin species: basis + ELL operator tables (the H_TBq of fermi_operator.h:42-242)."""

    def __init__(self, N, Q, ops=(), complex_tables=False):
        self.N, self.Q = N, Q
        self.slots = Z_SLOTS * (N if Q < 0 else Q) + 1
        dim = 1 << N
        self.state = np.zeros(dim, dtype=np.int32)
        self.address = np.zeros(dim, dtype=np.int32)
        self.count = int(lib().or_basis(N, Q, _ptr(self.state), _ptr(self.address)))
        self.state = self.state[: self.count].copy()
        self.complex = complex_tables
        n = self.count * self.slots
        self.Ax = np.zeros(n, dtype=np.complex128 if complex_tables else np.float64)
        self.Aj = np.zeros(n, dtype=np.int32)
        self.ops = []
        if ops:
            self.add(ops)

    def add(self, ops):
        """ops: iterable of (kind, i, j, value[, phase[, dag]]); dag (OP_FERMI1/2 only): bit 0 = 1
        for an annihilator on site i ('-'), bit 1 the same for site j."""
        ops = list(ops)
        self.ops += ops
        k = np.array([o[0] for o in ops], dtype=np.int32)
        i = np.array([o[1] for o in ops], dtype=np.int32)
        j = np.array([o[2] for o in ops], dtype=np.int32)
        v = np.array([o[3] for o in ops], dtype=np.float64)
        ph = np.array([o[4] if len(o) > 4 else 0.0 for o in ops], dtype=np.float64)
        dg = np.array([o[5] if len(o) > 5 else 0 for o in ops], dtype=np.int32)
        if self.complex:
            st = lib().or_tables_z(self.N, self.Q, _ptr(self.state), _ptr(self.address), self.count, len(ops),
                                   _ptr(k), _ptr(i), _ptr(j), _ptr(v), _ptr(ph), _ptr(self.Ax), _ptr(self.Aj),
                                   _ptr(dg))
        else:
            st = lib().or_tables(self.N, self.Q, _ptr(self.state), _ptr(self.address), self.count, len(ops),
                                 _ptr(k), _ptr(i), _ptr(j), _ptr(v), _ptr(self.Ax), _ptr(self.Aj), _ptr(dg))
        if st != 0:
            raise RuntimeError(f"oracle table build failed with status {st}")
        return self

    def dense(self):
        """H_TBq::Write (fermi_operator.h:227-235): M[j, Aj] += Ax over the nonzero slots, slot by
        slot in ascending order (numpy; each slot touches each row once)."""
        n = self.count
        M = np.zeros((n, n), dtype=self.Ax.dtype)
        rows = np.arange(n)
        for s_ in range(self.slots):
            ax = self.Ax[s_ * n:(s_ + 1) * n]
            aj = self.Aj[s_ * n:(s_ + 1) * n]
            nz = ax != 0
            np.add.at(M, (rows[nz], aj[nz]), ax[nz])
        return M


class Hubbard:
    """Oracle H = H_up (x) 1 + 1 (x) H_dn + U diag (HubMatrix_GPU.cuh:30-182)."""

    def __init__(self, up: Species, dn: Species, U=0.0, mask=None, Uvec=None):
        self.up, self.dn, self.U = up, dn, float(U)
        self.mask = (1 << up.N) - 1 if mask is None else int(mask)
        self.Uvec = None if Uvec is None else np.ascontiguousarray(Uvec, dtype=np.float64)
        self.cplx = bool(up.complex or dn.complex)
        self.dim_up, self.dim_dn = up.count, dn.count
        self.dim = self.dim_up * self.dim_dn
        AxUP = up.Ax.astype(np.complex128) if self.cplx else up.Ax
        AxDN = dn.Ax.astype(np.complex128) if self.cplx else dn.Ax
        self._keep = (AxUP, AxDN)
        self._op = lib().or_op_new(self.dim_up, self.dim_dn, up.slots, _ptr(AxUP), _ptr(up.Aj), dn.slots,
                                   _ptr(AxDN), _ptr(dn.Aj), _ptr(up.state), _ptr(dn.state), self.U, self.mask,
                                   _ptr(self.Uvec), int(self.cplx))

    def __del__(self):
        try:
            lib().or_op_free(self._op)
        except Exception:
            pass

    def dtype(self):
        return np.complex128 if self.cplx else np.float64

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=self.dtype())
        y = np.empty_like(x)
        lib().or_op_apply(self._op, _ptr(x), _ptr(y))
        return y

    def apply_rows(self, x, y, row0, nrows):
        """rows [row0, row0+nrows) of y = H x (real tables/vectors); y is written in place."""
        if self.cplx:
            raise ValueError("apply_rows is real-only")
        up, dn = self.up, self.dn
        lib().or_hubbard_apply_rows(self.dim_up, self.dim_dn, up.slots, _ptr(up.Ax), _ptr(up.Aj), dn.slots,
                                    _ptr(dn.Ax), _ptr(dn.Aj), _ptr(up.state), _ptr(dn.state), self.U, self.mask,
                                    _ptr(self.Uvec), _ptr(x), _ptr(y), int(row0), int(nrows))

    def dense(self):
        """Hubbard::Write (HubMatrix_GPU.cuh:184-230, kernel write_): H[r, c] with r = u*dim_dn + d;
        per row the nonzero UP slots in slot order (column AjUP*dim_dn + d), then the nonzero DN
        slots (column u*dim_dn + AjDN), then the interaction on the diagonal (numpy)."""
        up, dn = self.up, self.dn
        du, dd = self.dim_up, self.dim_dn
        D = du * dd
        H = np.zeros((D, D), dtype=self.dtype())
        u = np.repeat(np.arange(du), dd)
        d = np.tile(np.arange(dd), du)
        r = np.arange(D)
        for s_ in range(up.slots):
            ax = up.Ax[s_ * du:(s_ + 1) * du][u]
            aj = up.Aj[s_ * du:(s_ + 1) * du][u].astype(np.int64)
            nz = ax != 0
            np.add.at(H, (r[nz], (aj * dd + d)[nz]), ax[nz])
        for s_ in range(dn.slots):
            ax = dn.Ax[s_ * dd:(s_ + 1) * dd][d]
            aj = dn.Aj[s_ * dd:(s_ + 1) * dd][d].astype(np.int64)
            nz = ax != 0
            np.add.at(H, (r[nz], (u * dd + aj)[nz]), ax[nz])
        if self.U != 0.0 or self.Uvec is not None:
            dbl = up.state[u].astype(np.int64) & dn.state[d].astype(np.int64) & self.mask
            if self.Uvec is None:
                dg = self.U * np.array([bin(int(x)).count("1") for x in dbl], dtype=np.float64)
            else:
                dg = np.array([sum(self.Uvec[i] for i in range(up.N) if (int(x) >> i) & 1) for x in dbl])
            H[r, r] += dg
        return H

    def lanczos_eigen(self, f, maxN=1000, tol=None):
        """Lanczos<T>::eigen (HubMatrix_GPU.cuh:561-613).  Returns (status, E0, evec, a, b, nsteps)."""
        if tol is None:
            tol = 0.01 * np.sqrt(np.finfo(float).eps)
        f = np.ascontiguousarray(f, dtype=self.dtype()).copy()
        evec = np.empty_like(f)
        a = np.zeros(maxN + 1)
        b = np.zeros(maxN + 2)
        e0 = C.c_double(0.0)
        ns = C.c_int(0)
        st = lib().or_lanczos_eigen(self._op, maxN, tol, _ptr(f), C.byref(e0), _ptr(evec), _ptr(a), _ptr(b),
                                    C.byref(ns))
        n = ns.value
        return st, e0.value, evec, a[:n].copy(), b[: n + 1].copy(), n

    def lanczos_lowest_2vec(self, maxN=1000, tol=None, seed=235, verbose=False):
        """Eigenvalue-only Lanczos from the seed-235 start with two D-vectors (same recurrence and
        test as lanczos_eigen).  Returns (status, E0, a, b, nsteps)."""
        if tol is None:
            tol = 0.01 * np.sqrt(np.finfo(float).eps)
        a = np.zeros(maxN + 1)
        b = np.zeros(maxN + 2)
        e0 = C.c_double(0.0)
        ns = C.c_int(0)
        st = lib().or_lanczos_lowest_2vec(self._op, maxN, tol, seed, C.byref(e0), _ptr(a), _ptr(b), C.byref(ns),
                                          int(verbose))
        n = ns.value
        return st, e0.value, a[:n].copy(), b[: n + 1].copy(), n

    def lanczos_tridiagonal(self, f, maxN=200):
        """Lanczos<T>::tridiagonal (HubMatrix_GPU.cuh:541-559): a[0..maxN-1], b[0..maxN]."""
        f = np.ascontiguousarray(f, dtype=self.dtype())
        a = np.zeros(maxN)
        b = np.zeros(maxN + 1)
        lib().or_lanczos_tridiagonal(self._op, maxN, _ptr(f), _ptr(a), _ptr(b))
        return a, b


def s2_apply(up: Species, dn: Species, x):
    """TotalSpin kernel::Hubbard_mv_ (HubMatrix_TotalSpin.cuh:20-66)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    # full-length address tables (2^N), as the TotalSpin::Hubbard ctor uploads them
    aup = np.zeros(1 << up.N, dtype=np.int32)
    adn = np.zeros(1 << dn.N, dtype=np.int32)
    aup[up.state] = np.arange(up.count, dtype=np.int32)
    adn[dn.state] = np.arange(dn.count, dtype=np.int32)
    lib().or_s2_apply(up.N, up.count, dn.count, _ptr(up.state), _ptr(dn.state), _ptr(aup), _ptr(adn), _ptr(x),
                      _ptr(y))
    return y


def random_vector(n, seed=235, complex_=False, normalize=True):
    """GenRand/GenComplexRand + random_vector (HubMatrix_GPU.cuh:504-539, LanczosSolver.cuh:38-55)."""
    if complex_:
        v = np.empty(n, dtype=np.complex128)
        lib().or_random_vector_z(n, seed, _ptr(v))
        if normalize:
            lib().or_normalize_z(n, _ptr(v))
    else:
        v = np.empty(n, dtype=np.float64)
        lib().or_random_vector(n, seed, _ptr(v))
        if normalize:
            lib().or_normalize(n, _ptr(v))
    return v


def tridiag_lowest(d, e):
    d = np.ascontiguousarray(d, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64)
    z = np.zeros(len(d))
    lam = lib().or_tridiag_lowest(len(d), _ptr(d), _ptr(e), _ptr(z))
    return lam, z


def tridiag_eigvals(d, e):
    d = np.ascontiguousarray(d, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64)
    w = np.zeros(len(d))
    lib().or_tridiag_eigvals(len(d), _ptr(d), _ptr(e), _ptr(w))
    return w


def krylov_tridiagonal(H: Hubbard, f, max_steps=1000, btol=1e-10):
    """Krylov<C>::tridiagonal (HubMatrix_GPU.cuh:660-680), restated with numpy on the oracle H.v:
    f is used as given (complex128), cq = H f - b.back() cf_, a = Re<cq, f> (conj on the first
    argument, complex_prod at :325-329), cq -= a f, b = |cq|, stop after appending b < btol.
    Returns (a[n], b[n+1]) with b[0] = 0."""
    f = np.array(f, dtype=np.complex128)
    cf = np.zeros_like(f)
    a, b = [], [0.0]
    for _ in range(max_steps):
        q = -b[-1] * cf + H.apply(f)
        a.append(float(np.vdot(q, f).real))
        q = -a[-1] * f + q
        b.append(float(np.sqrt(np.vdot(q, q).real)))
        if b[-1] < btol:
            break
        cf = f
        f = (1.0 / b[-1]) * q
    return np.array(a), np.array(b)


def tridiag_eigen(a, b):
    """LAPACKf77::dsted(compz='I') as Krylov::Make_tridiagonal_eigen uses it
    (HubMatrix_GPU.cuh:683-692, MAGMA_LAPACK.h:296-326), here LAPACK's own tridiagonal solver
    through scipy: (m ascending, z column-major with z[j*n + i] = component i of vector j)."""
    from scipy.linalg import eigh_tridiagonal
    a = np.asarray(a, dtype=np.float64)
    n = len(a)
    m, V = eigh_tridiagonal(a, np.asarray(b, dtype=np.float64)[1:n])
    return m, np.ascontiguousarray(V.T).reshape(-1)


def krylov_time_evol(H: Hubbard, phi0, a, b, m, z, dt, phi=None):
    """Krylov<C>::Krylov_time_evol (HubMatrix_GPU.cuh:694-724): phi += sum_i Tri_i q_i with
    Tri_i = sum_j (cos(m_j dt) - i sin(m_j dt)) z[j n] z[j n + i], the q_i rebuilt by the same
    recurrence (q_{i+1} = ((H q_i - b_i q_{i-1}) - a_i q_i) / b_{i+1}).  Returns phi."""
    n = len(a)
    q = np.array(phi0, dtype=np.complex128)
    phi = np.zeros_like(q) if phi is None else np.array(phi, dtype=np.complex128)
    cf = np.zeros_like(q)
    zz = np.asarray(z).reshape(n, n)  # zz[j, i] = z[j n + i]
    for i in range(n):
        tri = np.sum((np.cos(m * dt) - 1j * np.sin(m * dt)) * zz[:, 0] * zz[:, i])
        phi = tri * q + phi
        if i + 1 == n:
            break
        cq = -b[i] * cf + H.apply(q)
        cq = -a[i] * q + cq
        cf = q
        q = (1.0 / b[i + 1]) * cq
    return phi


def continued_fraction(a, b, energy, norm, w):
    """correlator_continued_fraction loop (HubMatrix_GPU.cuh:829-834)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.complex128)
    g = np.empty_like(w)
    lib().or_continued_fraction(len(a), _ptr(a), _ptr(b), float(energy), float(norm), len(w), _ptr(w), _ptr(g))
    return g


def drude_combine(w, lam, kinetic, omega_step=0.004, omega_cutoff=0.01, vol=8.0 * np.sqrt(3.0)):
    """drude_weight_Jx Kubo combination (drude_v2_cal_Kagome_randomHop.cu:198-217)."""
    w = np.ascontiguousarray(w, dtype=np.complex128)
    lam = np.ascontiguousarray(lam, dtype=np.complex128)
    return lib().or_drude_combine(len(w), _ptr(w), _ptr(lam), float(kinetic), omega_step, omega_cutoff, vol)


def drude_weight(H: Hubbard, J: Hubbard, K: Hubbard, E0, psi, nomega=8000, omega_step=0.004, delta=0.01,
                 omega_cutoff=0.01, vol=8.0 * np.sqrt(3.0), nsteps=200):
    """drude_weight_Jx (drude_v2_cal_Kagome_randomHop.cu:183-220) via compute_sigma_omega
    (HubMatrix_GPU.cuh:837-849): phi = J psi; |phi|^2; 200-step Lanczos from phi/|phi|; continued
    fraction on the omega grid; D = (-<K> - 2 Vertex)/VOL."""
    phi = J.apply(psi)
    norm = float(np.vdot(phi, phi).real)
    phi = phi / np.sqrt(norm)
    a, b = H.lanczos_tridiagonal(phi, nsteps)
    w = omega_step * (np.arange(nomega) + 1.0) + 1j * delta
    lam = continued_fraction(a, b, E0, norm, w)
    kin = float(np.vdot(psi, K.apply(psi)).real)
    return drude_combine(w, lam, kin, omega_step, omega_cutoff, vol), kin, lam


def ted_time_evol(W, V, phi0, dt, phi=None):
    """tED<C>::tED_time_evol (HubMatrix_GPU.cuh:784-797): phi += sum_i e^{-i W_i dt} <V_i|phi0> V_i
    over the eigenpairs (W ascending, V[:, i] eigenvector i) of a dense eigensolver (numpy/LAPACK
    here, cuSOLVER syevd in the reference)."""
    phi0 = np.asarray(phi0, dtype=np.complex128)
    phi = np.zeros_like(phi0) if phi is None else np.array(phi, dtype=np.complex128)
    c = (np.cos(W * dt) - 1j * np.sin(W * dt)) * (V.conj().T @ phi0)
    return phi + V @ c


def prev_permutation(seq):
    """std::prev_permutation on a list in place (the standard algorithm: find the last i with
    seq[i-1] > seq[i], swap seq[i-1] with the rightmost element smaller than it, reverse the
    tail).  Returns False (and leaves the last permutation reversed to the largest) at the end."""
    n = len(seq)
    i = n - 1
    while i > 0 and seq[i - 1] <= seq[i]:
        i -= 1
    if i == 0:
        seq.reverse()
        return False
    j = n - 1
    while seq[j] >= seq[i - 1]:
        j -= 1
    seq[i - 1], seq[j] = seq[j], seq[i - 1]
    seq[i:] = seq[i:][::-1]
    return True


def lowmem_basis(N, Q):
    """LOW_MEM::H_TBq::Labeling (TotalSpin/include/fermi_operator_lowmem.h:50-64): occ[0..Q-1] true,
    then every prev_permutation of occ; state = sum occ[i] 2^i.  Pure Python (small N)."""
    occ = [True] * Q + [False] * (N - Q)
    out = []
    while True:
        out.append(sum(1 << i for i in range(N) if occ[i]))
        if not prev_permutation(occ):
            break
    return np.array(out, dtype=np.uint64)


def num_threads():
    return lib().or_num_threads()

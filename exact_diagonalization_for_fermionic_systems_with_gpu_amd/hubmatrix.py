"""Hubbard operator, Thrust_Lin BLAS-1, Lanczos and the continued-fraction correlator.

Mirror of MBL_test/include/HubMatrix_GPU.cuh with the reference's class and method
names (struct Hubbard :30-182, Thrust_Lin :291-471, Lanczos :477-613,
correlator_continued_fraction / compute_sigma_omega :820-849).  Vectors are torch
tensors on the GPU (float64 or complex128, length dimUP*dimDN, row-major
[i_up*dimDN + i_dn]) standing in for thrust::device_vector; all arithmetic runs in
libedhip.so.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from ._lib import LanczosCfg, call, dptr, stream_ptr, EdError, ED_ENOCONV
from .fermi_operator import H_TBq


def _is_complex(t: torch.Tensor) -> int:
    return 1 if t.is_complex() else 0


def _raw(t: torch.Tensor):
    return dptr(t)


class Hubbard:
    """Hubbard(HqUP, HqDN, U=0.0[, dilution]) (HubMatrix_GPU.cuh:46-95).

    Uvec (extension): per-site interaction U_i replacing U*popcount(up&dn&dilution)."""

    def __init__(self, HqUP: H_TBq, HqDN: H_TBq, U: float = 0.0, dilution: int | None = None, Uvec=None,
                 from_tables: bool = False):
        """from_tables=True builds from the device ELL tables (the reference's uploaded
        Ax/Aj, HubMatrix_GPU.cuh:63-69); otherwise from the two H_TBq call lists, which
        lets the library generate hop descriptors on the fly."""
        if HqUP.N != HqDN.N:
            raise ValueError("both species must live on the same N sites")
        self.N = HqUP.N
        self.dimUP, self.dimDN = HqUP.count, HqDN.count
        self.numcolsUP, self.numcolsDN = HqUP.slots, HqDN.slots
        self.U = float(U)
        self.is_diluted = dilution is not None
        self.dilution = ((1 << self.N) - 1) if dilution is None else int(dilution)
        self.complex_tables = bool(HqUP.complex_tables or HqDN.complex_tables)
        if HqUP.complex_tables != HqDN.complex_tables:
            raise ValueError("both species must use the same table type")
        uv = None
        if Uvec is not None:
            uv = np.ascontiguousarray(Uvec, dtype=np.float64)
            if uv.shape != (self.N,):
                raise ValueError("Uvec needs N entries")
        self._uv = uv
        uvp = None if uv is None else uv.ctypes.data_as(C.c_void_p)
        h = C.c_void_p()
        if from_tables or HqUP.full or HqDN.full:
            AxU, AjU = HqUP.raw_tables()
            AxD, AjD = HqDN.raw_tables()
            call("ed_hubbard_create", self.N, self.dimUP, HqUP.slots, dptr(AxU), dptr(AjU), dptr(HqUP.state),
                 self.dimDN, HqDN.slots, dptr(AxD), dptr(AjD), dptr(HqDN.state), int(self.complex_tables), self.U,
                 self.dilution, uvp, C.byref(h))
        else:
            au, nu = HqUP.ops_array()
            ad, nd = HqDN.ops_array()
            call("ed_hubbard_create_ops", self.N, HqUP.Q, au, nu, HqDN.Q, ad, nd, int(self.complex_tables), self.U,
                 self.dilution, uvp, C.byref(h))
        self._h = h
        self.dim = self.dimUP * self.dimDN

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                call("ed_hubbard_destroy", self._h)
                self._h = None
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def Hubbard_mv(self, x: torch.Tensor, y: torch.Tensor):
        """y = H x (HubMatrix_GPU.cuh:163-182); y is overwritten.  Asynchronous."""
        self._check(x, y)
        call("ed_hubbard_apply", self._h, _is_complex(x), _raw(x), _raw(y), stream_ptr())

    def autotune(self, x: torch.Tensor, y: torch.Tensor, reps: int = 3) -> dict:
        """Time the H.v tile-shape candidates on x (y is overwritten) and keep the fastest for this
        operator; returns the choice.  Results do not depend on it."""
        self._check(x, y)
        from ._lib import HvTuning
        t = HvTuning()
        call("ed_hubbard_autotune", self.handle, _is_complex(x), _raw(x), _raw(y), int(reps), C.byref(t),
             stream_ptr())
        return {"slab_width": t.slab_width, "dn_cached": t.dn_cached, "dn_panels": t.dn_panels,
                "dn_rows": t.dn_rows, "row_order": t.row_order, "up_table": t.up_table,
                "ms_per_apply": t.ms_per_apply}

    def apply(self, x: torch.Tensor) -> torch.Tensor:
        y = torch.empty_like(x)
        self.Hubbard_mv(x, y)
        return y

    # --- slab layout (the solver's internal vector layout; no reference counterpart) ----------
    def slab_layout(self, enable: bool = True) -> int:
        """Build (or free) the slab layout of this real operator; returns the layout's vector length
        (0 when freed).  Lanczos.eigen on real vectors then runs its recurrence on it."""
        n = C.c_int64()
        call("ed_hubbard_slab_layout", self._h, int(bool(enable)), C.byref(n))
        self.slab_len = n.value
        return self.slab_len

    def slab_len_query(self) -> int:
        """length of the slab layout's vectors, 0 when the operator has none (no side effect)"""
        n = C.c_int64()
        call("ed_hubbard_slab_layout", self._h, -1, C.byref(n))
        self.slab_len = n.value
        return self.slab_len

    def _slab_vec(self, *ts):
        """slab vectors: contiguous float64 CUDA tensors of exactly the handle's layout length"""
        n = self.slab_len_query()
        if n == 0:
            raise ValueError("the operator has no slab layout (slab_layout(True) builds it)")
        for t in ts:
            if t.numel() != n or t.dtype != torch.float64 or not t.is_contiguous() or not t.is_cuda:
                raise ValueError(f"slab vectors are contiguous float64 CUDA tensors of {n} elements")
        return n

    def _rowmajor_real(self, t):
        self._check(t)
        if t.dtype != torch.float64:
            raise ValueError("the slab layout holds real (float64) vectors")

    def to_slab(self, x: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """reference layout -> slab layout (padding zero); asynchronous"""
        self._rowmajor_real(x)
        if out is None:
            out = torch.empty(self.slab_len_query(), dtype=torch.float64, device=x.device)
        self._slab_vec(out)
        call("ed_slab_from_rowmajor", self._h, _raw(x), _raw(out), stream_ptr())
        return out

    def from_slab(self, xl: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """slab layout -> reference layout; asynchronous"""
        self._slab_vec(xl)
        if out is None:
            out = torch.empty(self.dim, dtype=torch.float64, device=xl.device)
        self._rowmajor_real(out)
        call("ed_slab_to_rowmajor", self._h, _raw(xl), _raw(out), stream_ptr())
        return out

    def apply_slab(self, xl: torch.Tensor, yl: torch.Tensor):
        """yl = H xl on the slab layout (asynchronous)"""
        self._slab_vec(xl, yl)
        if xl.data_ptr() == yl.data_ptr():
            raise ValueError("apply_slab: x and y must be different buffers")
        call("ed_hubbard_apply_slab", self._h, _raw(xl), _raw(yl), stream_ptr())

    def slab_autotune(self, xl: torch.Tensor, yl: torch.Tensor, reps: int = 3) -> dict:
        """Time the slab H.v launch shapes (tile rows x gathers in flight, then tile order) on xl
        (yl is overwritten) and keep the fastest in the layout; returns the choice.  Results do not
        depend on the choice."""
        self._slab_vec(xl, yl)
        if xl.data_ptr() == yl.data_ptr():
            raise ValueError("slab_autotune: x and y must be different buffers")
        from ._lib import SlabTuning
        t = SlabTuning()
        call("ed_hubbard_slab_autotune", self._h, _raw(xl), _raw(yl), int(reps), C.byref(t), stream_ptr())
        return {"tile_rows": t.tile_rows, "gathers_in_flight": t.gathers_in_flight, "tile_order": t.tile_order,
                "ms_per_apply": t.ms_per_apply, "candidate_ms": [round(v, 3) for v in t.candidate_ms]}

    def slab_set_shape(self, tile_rows: int = 32, gathers_in_flight: int = 8, tile_order: int = 8):
        """pin the slab H.v launch shape (e.g. a choice of slab_autotune from an earlier run)"""
        call("ed_hubbard_slab_set_shape", self._h, int(tile_rows), int(gathers_in_flight), int(tile_order))

    def slab_gather_ceiling(self, xl: torch.Tensor, pass_: int):
        """the line gathers of one slab H.v pass alone (0: UP over xl, 1: DN over the layout's
        transposed scratch), with the kernels' tiles and tile order but no arithmetic or streams:
        a measurement probe of the gather cost; asynchronous, writes nothing the caller sees"""
        self._slab_vec(xl)
        call("ed_hubbard_slab_gather_ceiling", self._h, int(pass_), _raw(xl), stream_ptr())

    def Write(self, Hmat: torch.Tensor):
        """Hubbard::Write (HubMatrix_GPU.cuh:219-230): ADDS H into the dense D x D matrix Hmat
        (row-major, Hmat[i_up*dimDN + i_dn, col]; float64, complex128 for complex tables)."""
        self._check_dense(Hmat)
        call("ed_hubbard_write_dense", self._h, _raw(Hmat), stream_ptr())

    Write_pinned = Write  # HubMatrix_GPU.cuh:232-248 (the target is any device-visible buffer)

    def initialize(self, Hmat: torch.Tensor):
        """Hubbard::initialize (HubMatrix_GPU.cuh:257-277): zero the dense matrix."""
        self._check_dense(Hmat)
        call("ed_dense_zero", self.dim, 1 if Hmat.is_complex() else 0, _raw(Hmat), stream_ptr())

    def _check_dense(self, Hmat):
        dt = torch.complex128 if self.complex_tables else torch.float64
        if Hmat.numel() != self.dim * self.dim or Hmat.dtype != dt or not Hmat.is_cuda or not Hmat.is_contiguous():
            raise ValueError(f"Hmat must be a contiguous CUDA {dt} tensor of (dimUP*dimDN)^2 elements")

    def _check(self, *ts):
        for t in ts:
            if t.numel() != self.dim or not t.is_cuda or not t.is_contiguous():
                raise ValueError(f"vector must be a contiguous CUDA tensor of {self.dim} elements")
            if t.dtype not in (torch.float64, torch.complex128):
                raise ValueError("vectors are float64 or complex128")
        if self.complex_tables and not ts[0].is_complex():
            raise ValueError("complex tables need complex128 vectors")


class Thrust_Lin:
    """Vector BLAS-1 of HubMatrix_GPU.cuh:291-471.  norm() returns the SQUARED norm as in
    the reference (:377-386); ax/axpy return new vectors (:388-449)."""

    def __init__(self, dimUP: int, dimDN: int):
        self.dimUP, self.dimDN = int(dimUP), int(dimDN)

    @staticmethod
    def thrust_Dotprod(v1: torch.Tensor, v2: torch.Tensor):
        out = (C.c_double * 2)()
        call("ed_dot", v1.numel(), _is_complex(v1), _raw(v1), _raw(v2), out, stream_ptr())
        if v1.is_complex():
            return complex(out[0], out[1])
        return out[0]

    @staticmethod
    def norm(v: torch.Tensor) -> float:
        out = (C.c_double * 2)()
        call("ed_dot", v.numel(), _is_complex(v), _raw(v), _raw(v), out, stream_ptr())
        return out[0]

    @staticmethod
    def normalize(v: torch.Tensor) -> float:
        n2 = C.c_double()
        call("ed_normalize", v.numel(), _is_complex(v), _raw(v), C.byref(n2), stream_ptr())
        return n2.value

    @staticmethod
    def ax(s: float, v: torch.Tensor) -> torch.Tensor:
        w = v.clone()
        call("ed_scal", w.numel(), _is_complex(w), float(s), _raw(w), stream_ptr())
        return w

    @staticmethod
    def axpy(s: float, v1: torch.Tensor, v2: torch.Tensor) -> torch.Tensor:
        w = v2.clone()
        call("ed_axpy", w.numel(), _is_complex(w), float(s), _raw(v1), _raw(w), stream_ptr())
        return w


_REORTH = {False: 0, True: 1, 0: 0, 1: 1, 2: 2, "none": 0, "dgks": 1, "selective": 2}


def lanczos_orthogonality(a, b) -> np.ndarray:
    """omega[j] = estimated max_{k<j} |v_j . v_k| of a Lanczos run without reorthogonalisation,
    from its a[0..n-1] and b[0..n-1] (b[0] the start norm): the omega recurrence (Simon 1984),
    ed_lanczos_orthogonality.  Host only."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    n = len(a)
    if n < 1 or len(b) < n:
        raise ValueError("lanczos_orthogonality: need len(b) >= len(a) >= 1")
    out = np.zeros(n)
    P = C.POINTER(C.c_double)
    call("ed_lanczos_orthogonality", n, a.ctypes.data_as(P), b.ctypes.data_as(P), out.ctypes.data_as(P))
    return out


class Lanczos(Thrust_Lin):
    """Lanczos<T> (HubMatrix_GPU.cuh:477-613).  a, b are host lists; b starts as [0.0].
    reorth: 0 / False (the reference's recurrence), 1 / True / "dgks" (a second Gram-Schmidt pass
    of f against v_N), 2 / "selective" (Parlett-Scott selective orthogonalisation against the
    converged Ritz vectors, the Lanczos basis kept on the device)."""

    def __init__(self, dimUP: int, dimDN: int, max_steps: int = 1000, tol: float | None = None,
                 check_every: int = 10, reorth=False):
        super().__init__(dimUP, dimDN)
        self.a: list[float] = []
        self.b: list[float] = [0.0]
        self.max_steps = int(max_steps)
        self.tol = 0.01 * math.sqrt(np.finfo(float).eps) if tol is None else float(tol)
        self.check_every = int(check_every)
        if reorth not in _REORTH:
            raise ValueError(f"reorth must be one of {sorted(map(str, _REORTH))}")
        self.reorth = _REORTH[reorth]
        self.nsteps = 0
        self.converged = False

    def clear(self):
        """HubMatrix_GPU.cuh:523-529"""
        self.a = []
        self.b = [0.0]

    def random_vector(self, f: torch.Tensor, seed: int = 235):
        """GenRand (seed 235) + normalize (HubMatrix_GPU.cuh:504-539)"""
        call("ed_random_vector", f.numel(), seed, _is_complex(f), _raw(f), 1, stream_ptr())

    def eigen(self, Hub: Hubbard, f: torch.Tensor, evec: torch.Tensor | None) -> float:
        """Lanczos<T>::eigen (HubMatrix_GPU.cuh:561-613): returns eval, fills evec (None: the
        eigenvalue only, no Ritz rebuild).  f is clobbered.  Raises EdError(ED_ENOCONV) if the
        convergence test is not met (evec/eval then hold the last Ritz pair)."""
        if evec is None:
            Hub._check(f)
        else:
            Hub._check(f, evec)
        cfg = LanczosCfg(self.max_steps, self.tol, self.check_every, int(self.reorth))
        ev = C.c_double()
        a = (C.c_double * (self.max_steps + 2))()
        b = (C.c_double * (self.max_steps + 3))()
        ns = C.c_int()
        from ._lib import lib
        st = lib().ed_lanczos_eigen(Hub.handle, _is_complex(f), C.byref(cfg), _raw(f), C.byref(ev), _raw(evec),
                                    a, b, C.byref(ns), stream_ptr())
        n = ns.value
        self.a = [a[k] for k in range(n)]
        self.b = [b[k] for k in range(n + 1)]
        self.nsteps = n
        self.converged = st == 0
        self.eval = ev.value
        if st not in (0, ED_ENOCONV):
            from ._lib import check
            check(st)
        if st == ED_ENOCONV:
            raise EdError(st, f"Lanczos convergence test not met in {self.max_steps} steps (E0={ev.value})")
        return ev.value

    def lowest(self, Hub: Hubbard, f: torch.Tensor, nev: int, evec: torch.Tensor | None = None) -> np.ndarray:
        """The nev lowest eigenvalues (ascending) by one Lanczos run with selective
        orthogonalisation (ed_lanczos_lowest: the SymEigsSolver "SmallestAlge" role,
        TotalSpin/include/SymEigsSolver.cuh:332-359); evec (optional) gets the lowest one's Ritz
        vector.  f is clobbered.  Raises EdError(ED_ENOCONV) if the test is not met for all nev."""
        if evec is None:
            Hub._check(f)
        else:
            Hub._check(f, evec)
        if not 1 <= nev <= 64:
            raise ValueError("lowest: nev must be 1..64")
        cfg = LanczosCfg(self.max_steps, self.tol, self.check_every, 2)
        ev = (C.c_double * nev)()
        ns = C.c_int()
        from ._lib import lib
        st = lib().ed_lanczos_lowest(Hub.handle, _is_complex(f), C.byref(cfg), _raw(f), int(nev), ev, _raw(evec),
                                     C.byref(ns), stream_ptr())
        self.nsteps = ns.value
        self.converged = st == 0
        vals = np.array(ev[:])
        self.eval = float(vals[0])
        if st not in (0, ED_ENOCONV):
            from ._lib import check
            check(st)
        if st == ED_ENOCONV:
            raise EdError(st, f"Lanczos test of the {nev} lowest not met in {self.max_steps} steps ({vals})")
        return vals

    def so_stats(self) -> dict:
        """selective-orthogonalisation statistics of this thread's last eigen() (ed_lanczos_so_stats)"""
        g, an, pu, om = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        call("ed_lanczos_so_stats", C.byref(g), C.byref(an), C.byref(pu), C.byref(om))
        return {"good_ritz_vectors": g.value, "analyses": an.value, "purges": pu.value, "omega_max": om.value}

    def tridiagonal(self, Hub: Hubbard, f: torch.Tensor, maxN: int = 200):
        """Lanczos<T>::tridiagonal (HubMatrix_GPU.cuh:541-559): appends maxN a's and b's."""
        Hub._check(f)
        a = (C.c_double * maxN)()
        b = (C.c_double * (maxN + 1))()
        call("ed_lanczos_tridiagonal", Hub.handle, _is_complex(f), maxN, _raw(f), a, b, stream_ptr())
        self.b[0] = b[0]
        self.a += [a[k] for k in range(maxN)]
        self.b += [b[k] for k in range(1, maxN + 1)]


def tridiag_eigen(a, b):
    """All eigenpairs of the tridiagonal T (diagonal a[n], off-diagonal b[1..n-1] of the
    reference's b list), as LAPACKf77::dsted(compz='I') returns them (MAGMA_LAPACK.h:296-326):
    (m ascending [n], z [n*n] column-major, z[j*n + i] = component i of eigenvector j)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = len(a)
    e = np.ascontiguousarray(np.asarray(b, dtype=np.float64)[1:n], dtype=np.float64)
    if len(e) == 0:
        e = np.zeros(1)
    m = np.empty(n)
    z = np.empty(n * n)
    call("ed_tridiag_eigen", n, a.ctypes.data_as(C.POINTER(C.c_double)), e.ctypes.data_as(C.POINTER(C.c_double)),
         m.ctypes.data_as(C.POINTER(C.c_double)), z.ctypes.data_as(C.POINTER(C.c_double)))
    return m, z


class Krylov(Thrust_Lin):
    """Krylov<C> (HubMatrix_GPU.cuh:619-728): real-time evolution psi(dt) = exp(-i H dt) psi0 in
    the Lanczos space of psi0 (complex128 vectors).  a, b (b starts as [0.0]), m, z are host
    arrays as in the reference."""

    def __init__(self, dimUP: int, dimDN: int, max_steps: int = 1000, btol: float = 1e-10):
        super().__init__(dimUP, dimDN)
        self.max_steps = int(max_steps)
        self.btol = float(btol)
        self.clear()

    def clear(self):
        """HubMatrix_GPU.cuh:730-739"""
        self.a: list[float] = []
        self.b: list[float] = [0.0]
        self.m = np.zeros(0)
        self.z = np.zeros(0)

    def tridiagonal(self, Hub: Hubbard, f: torch.Tensor):
        """Krylov<C>::tridiagonal (HubMatrix_GPU.cuh:660-680): recurrence from f as it is (no
        normalisation), at most max_steps steps, stopping after the first b < btol."""
        Hub._check(f)
        if not f.is_complex():
            raise TypeError("Krylov works on complex128 vectors (thrust::complex<C> in the reference)")
        a = (C.c_double * self.max_steps)()
        b = (C.c_double * (self.max_steps + 1))()
        n = C.c_int()
        call("ed_krylov_tridiagonal", Hub.handle, _raw(f), self.max_steps, self.btol, a, b, C.byref(n),
             stream_ptr())
        self.a = [a[k] for k in range(n.value)]
        self.b = [0.0] + [b[k] for k in range(1, n.value + 1)]

    def Make_tridiagonal_eigen(self, Hub: Hubbard, phi0: torch.Tensor):
        """HubMatrix_GPU.cuh:683-692: tridiagonal, then all eigenpairs of T (m, z)."""
        self.tridiagonal(Hub, phi0)
        self.m, self.z = tridiag_eigen(self.a, self.b)

    def Krylov_time_evol(self, Hub: Hubbard, phi0: torch.Tensor, phi: torch.Tensor, dt: float):
        """HubMatrix_GPU.cuh:694-724: phi += sum_i c_i q_i with c_i = sum_j exp(-i m_j dt) z_0j z_ij
        (phi accumulates; pass zeros for exp(-i H dt) phi0).  phi0 is not modified."""
        Hub._check(phi0, phi)
        n = len(self.a)
        a = np.ascontiguousarray(self.a, dtype=np.float64)
        b = np.ascontiguousarray(self.b, dtype=np.float64)
        m = np.ascontiguousarray(self.m, dtype=np.float64)
        z = np.ascontiguousarray(self.z, dtype=np.float64)
        if len(m) != n or len(z) != n * n:
            raise ValueError("Krylov_time_evol needs Make_tridiagonal_eigen first")
        P = C.POINTER(C.c_double)
        call("ed_krylov_time_evol", Hub.handle, _raw(phi0), n, a.ctypes.data_as(P), b.ctypes.data_as(P),
             m.ctypes.data_as(P), z.ctypes.data_as(P), float(dt), _raw(phi), stream_ptr())


def cusolverDsyevd(A: torch.Tensor, W: torch.Tensor):
    """cusolverDsyevd / cusolverZheevd (EIGEN_cusolver.h:16-63, 65-112) on rocSOLVER: all
    eigenpairs of the dense Hermitian matrix A (n x n, written by Hubbard.Write or H_TBq.Write).
    W (float64, n) receives the eigenvalues ascending, A the eigenvectors: A.view(n, n)[k] is the
    unit eigenvector of W[k] (the reference reads eigenvector k at A + k*n)."""
    n = int(round(math.sqrt(A.numel())))
    if n * n != A.numel() or W.numel() != n or W.dtype != torch.float64 or not A.is_contiguous():
        raise ValueError("A must be a contiguous n x n matrix and W a float64 vector of n")
    if A.dtype not in (torch.float64, torch.complex128):
        raise ValueError("A is float64 or complex128")
    call("ed_dense_eigh", n, 1 if A.is_complex() else 0, _raw(A), _raw(W), stream_ptr())


cusolverZheevd = cusolverDsyevd


class tED(Thrust_Lin):
    """tED<C> (HubMatrix_GPU.cuh:730-814): exact propagation in the eigenbasis of the dense
    Hamiltonian.  W (eigenvalues) and V (eigenvectors, V.view(D, D)[k] = eigenvector k) live on
    the device."""

    def __init__(self, dimUP: int, dimDN: int):
        super().__init__(dimUP, dimDN)
        self.dim = int(dimUP) * int(dimDN)
        self.W = None
        self.V = None

    def Make_eigen(self, Hmat: torch.Tensor):
        """HubMatrix_GPU.cuh:756-767: V = Hmat (copied), then the dense eigensolver."""
        if Hmat.numel() != self.dim * self.dim:
            raise ValueError("Hmat must hold (dimUP*dimDN)^2 elements")
        self.V = Hmat.reshape(-1).clone()
        self.W = torch.empty(self.dim, dtype=torch.float64, device=Hmat.device)
        cusolverDsyevd(self.V, self.W)

    def tED_time_evol(self, phi0: torch.Tensor, phi: torch.Tensor, dt: float):
        """HubMatrix_GPU.cuh:777-790: phi += sum_k exp(-i W_k dt) <V_k|phi0> V_k (phi accumulates,
        as in the reference; pass zeros for exp(-i H dt) phi0).  complex128 vectors."""
        if self.V is None:
            raise ValueError("tED_time_evol needs Make_eigen first")
        for t in (phi0, phi):
            if t.numel() != self.dim or t.dtype != torch.complex128 or not t.is_contiguous() or not t.is_cuda:
                raise ValueError("phi0 and phi are contiguous complex128 CUDA vectors of dimUP*dimDN")
        call("ed_ted_time_evol", self.dim, 1 if self.V.is_complex() else 0, _raw(self.V), _raw(self.W),
             _raw(phi0), float(dt), _raw(phi), stream_ptr())

    def clear(self):
        """HubMatrix_GPU.cuh:807-812"""
        self.W = None
        self.V = None


def correlator_continued_fraction(lanctri: Lanczos, H: Hubbard, nw: int, w, energy: float, phi: torch.Tensor,
                                  maxN: int = 200):
    """HubMatrix_GPU.cuh:820-835: g(w) = |phi|^2/(w+E-a0-b1^2/(w+E-a1-...)).  phi is
    normalized in place (as in the reference).  Returns g as a complex numpy array."""
    norm = lanctri.norm(phi)
    lanctri.normalize(phi)
    lanctri.tridiagonal(H, phi, maxN)
    a = np.ascontiguousarray(lanctri.a, dtype=np.float64)
    b = np.ascontiguousarray(lanctri.b, dtype=np.float64)
    w = np.ascontiguousarray(np.asarray(w, dtype=np.complex128)[:nw])
    g = np.empty_like(w)
    call("ed_continued_fraction", len(a), a.ctypes.data_as(C.POINTER(C.c_double)),
         b.ctypes.data_as(C.POINTER(C.c_double)), float(energy), float(norm),
         len(w), w.ctypes.data_as(C.POINTER(C.c_double)), g.ctypes.data_as(C.POINTER(C.c_double)))
    return g


def compute_sigma_omega(lanctri: Lanczos, H: Hubbard, O: Hubbard, E: float, evec: torch.Tensor, w,
                        maxN: int = 200):
    """HubMatrix_GPU.cuh:837-849: Lambda(w) for the correlator of O on the ground state."""
    lanctri.clear()
    phi = torch.empty_like(evec)
    O.Hubbard_mv(evec, phi)
    return correlator_continued_fraction(lanctri, H, len(w), w, E, phi, maxN)


# constants of the Drude drivers (drude_v2_cal_Kagome_randomHop.cu:14-21)
NOMEGA = 8000
OMEGA_STEP = 0.004
DELTA = 0.01
OMEGA_CUTOFF = 0.01


def drude_weight_Jx(lanc: Lanczos, Hmat: Hubbard, Jxmat: Hubbard, Kxmat: Hubbard, eval_: float,
                    evec: torch.Tensor, vol: float, nomega: int = NOMEGA, omega_step: float = OMEGA_STEP,
                    delta: float = DELTA, omega_cutoff: float = OMEGA_CUTOFF):
    """drude_weight_Jx (drude_v2_cal_Kagome_randomHop.cu:183-220): D = (-<K_x> - 2 Vertex)/VOL with
    sigma(w) = -Im Lambda(w)/(pi w) from the 200-step continued fraction and
    Vertex = sum_{w > cutoff} sigma(w) * step.  Returns (D, <K_x>, Lambda)."""
    omega = omega_step * (np.arange(nomega, dtype=np.float64) + 1.0) + 1j * delta
    Lam = compute_sigma_omega(lanc, Hmat, Jxmat, eval_, evec, omega)
    sigma = -1.0 / (math.pi * omega.real) * Lam.imag
    vertex = float(np.sum(np.where(omega.real > omega_cutoff, sigma * omega_step, 0.0)))
    KV = torch.empty_like(evec)
    Kxmat.Hubbard_mv(evec, KV)
    kinetic = lanc.thrust_Dotprod(evec, KV)
    kinetic = kinetic.real if isinstance(kinetic, complex) else kinetic
    return (-kinetic - 2.0 * vertex) / vol, kinetic, Lam

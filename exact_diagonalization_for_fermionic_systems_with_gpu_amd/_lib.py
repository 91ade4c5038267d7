"""ctypes binding of libedhip.so (include/edhip.h).

The library is built in-tree by build.py.  Loading fails loudly when it is missing:
there is no CPU fallback behind any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libedhip.so")

ED_OK, ED_EINVAL, ED_ENOMEM, ED_EBREAKDOWN, ED_ENOCONV, ED_EHIP, ED_EUNSUPPORTED, ED_ESLOT = range(8)
ED_OP_HOP, ED_OP_CURRENT, ED_OP_DENSITY, ED_OP_DENSITY2, ED_OP_HOP_PHASE, ED_OP_FERMI1, ED_OP_FERMI2 = range(7)

_p = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int
_d = C.c_double
_pd = C.POINTER(C.c_double)
_pi64 = C.POINTER(C.c_int64)
_pi = C.POINTER(C.c_int)


class EdOp(C.Structure):
    _fields_ = [("kind", C.c_int32), ("i", C.c_int32), ("j", C.c_int32), ("dag", C.c_int32),
                ("value", C.c_double), ("phase", C.c_double)]


class LanczosCfg(C.Structure):
    _fields_ = [("max_steps", C.c_int), ("tol", C.c_double), ("check_every", C.c_int), ("reorth", C.c_int)]


class HvTuning(C.Structure):
    _fields_ = [("slab_width", C.c_int), ("dn_cached", C.c_int), ("dn_panels", C.c_int), ("dn_rows", C.c_int),
                ("row_order", C.c_int), ("up_table", C.c_int), ("ms_per_apply", C.c_double)]


class SlabTuning(C.Structure):
    _fields_ = [("tile_rows", C.c_int), ("gathers_in_flight", C.c_int), ("tile_order", C.c_int),
                ("ms_per_apply", C.c_double), ("candidate_ms", C.c_double * 8)]


class BatchTuning(C.Structure):
    _fields_ = [("dn_rows", C.c_int), ("dn_panels", C.c_int), ("up_rows", C.c_int), ("up_panels", C.c_int),
                ("steps", C.c_int), ("seconds", C.c_double), ("candidate_s", C.c_double * 18)]


class EdError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"libedhip status {status}: {msg}")
        self.status = status


# name -> (restype, argtypes); every symbol declared in include/edhip.h
SIGNATURES = {
    "ed_last_error": (C.c_char_p, []),
    "ed_version": (_i32, []),
    "ed_device_sync": (_i32, []),
    "ed_basis_count": (_i32, [_i32, _i32, _pi64]),
    "ed_basis_build": (_i32, [_i32, _i32, _p, _p, _p]),
    "ed_basis_build_lowmem": (_i32, [_i32, _i32, _p, _p]),
    "ed_tables_build": (_i32, [_i32, _i32, _p, _p, _i64, _p, _i32, _i32, _p, _p, _p]),
    "ed_hubbard_create": (_i32, [_i32, _i64, _i32, _p, _p, _p, _i64, _i32, _p, _p, _p, _i32, _d, _i64, _p,
                                 C.POINTER(_p)]),
    "ed_hubbard_create_ops": (_i32, [_i32, _i32, _p, _i32, _i32, _p, _i32, _i32, _d, _i64, _p, C.POINTER(_p)]),
    "ed_hubbard_destroy": (_i32, [_p]),
    "ed_hubbard_info": (_i32, [_p, _pi64, _pi64, _pi]),
    "ed_hubbard_apply": (_i32, [_p, _i32, _p, _p, _p]),
    "ed_hubbard_autotune": (_i32, [_p, _i32, _p, _p, _i32, C.POINTER(HvTuning), _p]),
    "ed_hubbard_slab_layout": (_i32, [_p, _i32, _pi64]),
    "ed_slab_from_rowmajor": (_i32, [_p, _p, _p, _p]),
    "ed_slab_to_rowmajor": (_i32, [_p, _p, _p, _p]),
    "ed_hubbard_apply_slab": (_i32, [_p, _p, _p, _p]),
    "ed_hubbard_slab_gather_ceiling": (_i32, [_p, _i32, _p, _p]),
    "ed_hubbard_slab_set_shape": (_i32, [_p, _i32, _i32, _i32]),
    "ed_memory_stats": (_i32, [_pi64, _pi64, _i32]),
    "ed_slab_auto_decision": (_i32, [C.c_double, C.c_double, _i64, _i64, _i32]),
    "ed_hubbard_slab_autotune": (_i32, [_p, _p, _p, _i32, C.POINTER(SlabTuning), _p]),
    "ed_dense_zero": (_i32, [_i64, _i32, _p, _p]),
    "ed_hubbard_write_dense": (_i32, [_p, _p, _p]),
    "ed_tables_write_dense": (_i32, [_i64, _i32, _i32, _p, _p, _p, _p]),
    "ed_dense_eigh": (_i32, [_i64, _i32, _p, _p, _p]),
    "ed_ted_time_evol": (_i32, [_i64, _i32, _p, _p, _p, _d, _p, _p]),
    "ed_random_vector": (_i32, [_i64, C.c_uint32, _i32, _p, _i32, _p]),
    "ed_dot": (_i32, [_i64, _i32, _p, _p, _pd, _p]),
    "ed_dot_noalloc": (_i32, [_i64, _i32, _p, _p, _p, _p]),
    "ed_axpy": (_i32, [_i64, _i32, _d, _p, _p, _p]),
    "ed_scal": (_i32, [_i64, _i32, _d, _p, _p]),
    "ed_normalize": (_i32, [_i64, _i32, _p, _pd, _p]),
    "ed_lanczos_eigen": (_i32, [_p, _i32, C.POINTER(LanczosCfg), _p, _pd, _p, _pd, _pd, _pi, _p]),
    "ed_lanczos_lowest": (_i32, [_p, _i32, C.POINTER(LanczosCfg), _p, _i32, _pd, _p, _pi, _p]),
    "ed_lanczos_tridiagonal": (_i32, [_p, _i32, _i32, _p, _pd, _pd, _p]),
    "ed_tridiag_eigen": (_i32, [_i32, _pd, _pd, _pd, _pd]),
    "ed_tridiag_lanczos_lowest": (_i32, [_i32, _pd, _pd, C.c_double, _i32, C.c_double, C.c_double, _i32, _pd, _pi, _pi]),
    "ed_lanczos_orthogonality": (_i32, [_i32, _pd, _pd, _pd]),
    "ed_lanczos_so_stats": (_i32, [_pi, _pi, _pi, _pd]),
    "ed_krylov_tridiagonal": (_i32, [_p, _p, _i32, _d, _pd, _pd, _pi, _p]),
    "ed_krylov_time_evol": (_i32, [_p, _p, _i32, _pd, _pd, _pd, _pd, _d, _p, _p]),
    "ed_tridiag_lowest": (_i32, [_i32, _pd, _pd, _pd, _pd]),
    "ed_tridiag_eigvals": (_i32, [_i32, _pd, _pd, _pd]),
    "ed_eigs_lowest": (_i32, [_p, _i32, _i32, _i32, _i32, _d, C.c_uint32, _pd, _p, _pi, _pi, _p]),
    "ed_hubbard_apply_host": (_i32, [_p, _i32, _p, _p, _p]),
    "ed_batch_create": (_i32, [_i32, _p, _i32, C.POINTER(_p)]),
    "ed_batch_destroy": (_i32, [_p]),
    "ed_batch_lowest": (_i32, [_p, C.POINTER(LanczosCfg), C.c_uint32, _i32, _pd, _pi, _pi, _p]),
    "ed_batch_lowest_k": (_i32, [_p, C.POINTER(LanczosCfg), _i32, C.c_uint32, _i32, _pd, _pi, _pi, _p, _p]),
    "ed_batch_set_shape": (_i32, [_p, _i32, _i32, _i32, _i32]),
    "ed_batch_autotune": (_i32, [_p, _i32, C.POINTER(BatchTuning), _p]),
    "ed_bond_bilinear": (_i32, [_p, _i32, _i32, _p, _p, _pd, _p]),
    "ed_site_double_occupancy": (_i32, [_p, _i32, _p, _p, _pd, _p]),
    "ed_site_density": (_i32, [_p, _i32, _p, _p, _pd, _p]),
    "ed_s2_create": (_i32, [_i32, _i64, _p, _p, _i64, _p, _p, C.POINTER(_p)]),
    "ed_s2_destroy": (_i32, [_p]),
    "ed_s2_apply": (_i32, [_p, _i32, _p, _p, _p]),
    "ed_continued_fraction": (_i32, [_i32, _pd, _pd, _d, _d, _i32, _pd, _pd]),
    "ed_entanglement_entropy": (_i32, [_i32, _i32, _i32, _i64, _i32, _p, _pd, _pi64, _p]),
    "ed_entanglement_entropy_vn": (_i32, [_i32, _i32, _i32, _i64, _i32, _p, _pd, _p]),
    "ed_entanglement_entropy_vn_batch": (_i32, [_i32, _i32, _i32, _i64, _i32, _i32, _p, _pd, _p]),
    "ed_entanglement_spectrum": (_i32, [_i32, _i32, _i32, _i64, _i32, _p, _pd, _i64, _pi64, _p]),
    "ed_entanglement_stats": (_i32, [_pd, _pd, _pd, _i32]),
    "ed_sym_eigvals": (_i32, [_i64, _p, _i64, _i32, _p, _p]),
}

_LIB = None


def lib():
    """Load libedhip.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               "`python -m exact_diagonalization_for_fermionic_systems_with_gpu_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(status):
    if status != ED_OK:
        msg = lib().ed_last_error()
        raise EdError(status, msg.decode() if msg else "")
    return status


def call(name, *args):
    return check(getattr(lib(), name)(*args))


def dptr(t):
    """device pointer of a torch tensor (or None)"""
    return None if t is None else C.c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def darr(a):
    import numpy as np
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_pd)

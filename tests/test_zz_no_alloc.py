"""ed_dot_noalloc, the allocation-free variant of ed_dot: through the C ABI it gives ed_dot's value
bit for bit (same kernels, same reduction order) on two calls in a row, the second on a smaller
input, for real and complex vectors; it rejects a size beyond its fixed partial buffer before any
launch."""
import ctypes as C

import numpy as np
import pytest
import torch

from exact_diagonalization_for_fermionic_systems_with_gpu_amd._lib import lib, stream_ptr

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _dot(n, cplx, a, b):
    out = (C.c_double * 2)()
    assert lib().ed_dot(n, cplx, _ptr(a), _ptr(b), out, stream_ptr()) == 0
    return out[0], (out[1] if cplx else 0.0)


def _dot_noalloc(n, cplx, a, b):
    d_out = torch.full((2,), float("nan"), dtype=torch.float64, device=DEV)
    assert lib().ed_dot_noalloc(n, cplx, _ptr(a), _ptr(b), _ptr(d_out), stream_ptr()) == 0
    torch.cuda.synchronize()
    return tuple(d_out.cpu().numpy())


@pytest.mark.parametrize("cplx", [0, 1])
def test_dot_noalloc_matches_dot(cplx):
    print(f"ed_dot_noalloc cplx={cplx}", flush=True)
    g = torch.Generator(device=DEV).manual_seed(11 + cplx)
    for n in (3_000_017, 1_234_567):  # second call smaller: no stale partials may leak in
        m = 2 * n if cplx else n
        a = torch.randn(m, dtype=torch.float64, device=DEV, generator=g)
        b = torch.randn(m, dtype=torch.float64, device=DEV, generator=g)
        ref = _dot(n, cplx, a, b)
        got = _dot_noalloc(n, cplx, a, b)
        assert np.array_equal(np.array(got).view(np.uint64), np.array(ref).view(np.uint64)), (n, got, ref)
        if cplx:
            za = a.view(torch.complex128)
            zb = b.view(torch.complex128)
            z = torch.sum(torch.conj(za) * zb).item()
            assert abs(got[0] - z.real) <= 1e-9 * abs(z) and abs(got[1] - z.imag) <= 1e-9 * abs(z)
        else:
            assert got[1] == 0.0
            assert abs(got[0] - float(torch.dot(a, b))) <= 1e-9 * float(torch.linalg.vector_norm(a) ** 2)


def test_dot_noalloc_rejects_oversize():
    a = torch.zeros(16, dtype=torch.float64, device=DEV)
    d_out = torch.zeros(2, dtype=torch.float64, device=DEV)
    n = (1 << 32) + 1  # one more than the fixed buffer holds; checked before any launch
    st = lib().ed_dot_noalloc(n, 0, _ptr(a), _ptr(a), _ptr(d_out), stream_ptr())
    assert st != 0
    assert b"exceed" in lib().ed_last_error()

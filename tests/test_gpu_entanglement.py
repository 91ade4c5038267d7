"""GPU entanglement entropy (per-block rho_A with FP64 MFMA + rocSOLVER) against the oracle:
the reference's dense first-appearance matrix + SVD on small systems, and the block oracle on
the 12-site Kagome ground state (6|6 sites).  Bar: 1e-8 relative on S."""
import math

import numpy as np
import pytest
import torch

from oracle import ee_oracle as E
from oracle import oracle as O
from exact_diagonalization_for_fermionic_systems_with_gpu_amd import H_TBq, Hubbard, Lanczos
from exact_diagonalization_for_fermionic_systems_with_gpu_amd import lattices as L
from exact_diagonalization_for_fermionic_systems_with_gpu_amd.entangle import EEmatHubAB_GPU

pytestmark = pytest.mark.gpu


def _state(n, seed, cplx):
    r = np.random.default_rng(seed)
    v = r.normal(size=n) + (1j * r.normal(size=n) if cplx else 0)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("N,Qu,Qd,A,cplx", [(6, 3, 3, [0, 1, 2], False), (8, 4, 3, [0, 1, 2, 3], True),
                                            (7, 3, 3, [1, 2, 5], False), (10, 5, 5, [0, 1, 2, 3, 4], False),
                                            (10, 4, 6, [0, 3, 4, 7, 9], True)])
def test_entropy_vs_reference_matrix(N, Qu, Qd, A, cplx):
    print(f"ee N={N} {Qu}/{Qd} A={A} cplx={cplx}", flush=True)
    hu, hd = H_TBq(N, Qu), H_TBq(N, Qd)
    psi = _state(hu.count * hd.count, N + Qu, cplx)
    sref, s = E.reference_entropy(N, Qu, Qd, A, psi)
    lam_o, sig_o = E.block_spectra(N, Qu, Qd, sum(1 << a for a in A), psi)
    ee = EEmatHubAB_GPU(A, len(A), hu, hd)
    svn, sr, nsv = ee.entropy(torch.from_numpy(psi).cuda())
    assert abs(sr - sref) <= 1e-10 * max(1.0, abs(sref))
    assert abs(svn - E.entropy_from_lambda(lam_o)) <= 1e-10 * max(1.0, abs(svn))
    lam = ee.spectrum(torch.from_numpy(psi).cuda())
    assert abs(lam.sum() - 1.0) < 1e-12
    k = min(len(lam), len(lam_o))
    assert np.allclose(lam[:k], np.sort(lam_o)[::-1][:k], atol=1e-13)


def test_gram_mfma_large_block():
    # one block of 300 x 900 (N=12: A = 6 sites, n_A^up = 3, n_A^dn = 3 -> 400 x 400 ...) via a random state
    N, Qu, Qd = 12, 6, 6
    hu = H_TBq(N, Qu)
    psi = _state(hu.count * hu.count, 5, False)
    ee = EEmatHubAB_GPU(list(range(6)), 6, hu, hu)
    lam = ee.spectrum(torch.from_numpy(psi).cuda())
    lam_o, _ = E.block_spectra(N, Qu, Qd, 0b111111, psi)
    assert np.allclose(lam, np.sort(lam_o)[::-1], atol=1e-13)


def test_kagome12_ground_state_entropy():
    ops = L.hubbard_ops(L.KAGOME12, t=-1.0)
    hq = H_TBq(12, 6)
    hq.add_ops(ops)
    H = Hubbard(hq, hq, 4.0)
    f = torch.empty(H.dim, dtype=torch.float64, device="cuda")
    ev = torch.empty_like(f)
    lz = Lanczos(hq.count, hq.count, tol=1e-12, max_steps=300)
    lz.random_vector(f)
    lz.eigen(H, f, ev)
    psi = ev.cpu().numpy()
    ee = EEmatHubAB_GPU(list(range(6)), 6, hq, hq)
    svn, sr, nsv = ee.entropy(ev)
    lam_o, sig_o = E.block_spectra(12, 6, 6, 0b111111, psi)
    assert abs(svn - E.entropy_from_lambda(lam_o)) <= 1e-8 * abs(svn)
    assert abs(sr - E.entropy_from_sigma(sig_o)) <= 1e-8 * abs(sr)


def _entropy_vn(N, Q, nA, psi, cplx=False):
    import ctypes as C
    from exact_diagonalization_for_fermionic_systems_with_gpu_amd._lib import call, dptr, stream_ptr
    out = np.zeros(1)
    call("ed_entanglement_entropy_vn", N, Q, Q, (1 << nA) - 1, int(cplx), dptr(psi), out.ctypes.data_as(C.POINTER(C.c_double)),
         stream_ptr())
    return float(out[0])


@pytest.mark.parametrize("N,Q,nA,cplx", [(12, 6, 6, False), (12, 6, 6, True), (16, 8, 8, False)])
def test_concurrent_block_solves_match_sequential(N, Q, nA, cplx, monkeypatch):
    """the blocks solved on four worker streams give the one-stream spectra bit for bit (same
    arithmetic per block; the threshold ED_EE_PARALLEL_MIN selects the path).  16 sites 8|8 has
    blocks of 4,900 rows, which take the concurrent path by default.  The Schmidt values (SVD of
    each block) are compared at 12 sites only: rocSOLVER's bidiagonal SVD takes minutes on the
    4,900-row blocks."""
    print(f"ee concurrent vs sequential N={N} {Q}/{Q} A=0..{nA - 1} cplx={cplx}", flush=True)
    hq = H_TBq(N, Q)
    psi = torch.from_numpy(_state(hq.count * hq.count, N, cplx)).cuda()
    ee = EEmatHubAB_GPU(list(range(nA)), nA, hq, hq)
    runs = {}
    for label, par_min in (("one stream", "1000000000"), ("four streams", "1")):
        monkeypatch.setenv("ED_EE_PARALLEL_MIN", par_min)
        print(f"  {label}: spectrum", flush=True)
        lam = ee.spectrum(psi)
        print(f"  {label}: S_vN", flush=True)
        svn = _entropy_vn(N, Q, nA, psi, cplx)
        full = ee.entropy(psi) if N <= 12 else None
        runs[label] = (lam, svn, full)
    (lam_1, s_1, f_1), (lam_4, s_4, f_4) = runs["one stream"], runs["four streams"]
    assert np.array_equal(lam_1, lam_4)
    assert s_1 == s_4 and f_1 == f_4
    assert abs(lam_4.sum() - 1.0) < 1e-12
    assert abs(s_4 - float(-np.sum(lam_4[lam_4 > 0] * np.log(lam_4[lam_4 > 0])))) <= 1e-10 * s_4
    if N == 12:
        lam_o, _ = E.block_spectra(N, Q, Q, (1 << nA) - 1, psi.cpu().numpy())
        assert np.allclose(lam_4, np.sort(lam_o)[::-1], atol=1e-13)



_FIRST_CALL_SCRIPT = r"""
import ctypes as C, os, sys
sys.path.insert(0, sys.argv[1])
import numpy as np, torch
from exact_diagonalization_for_fermionic_systems_with_gpu_amd import Lanczos
from exact_diagonalization_for_fermionic_systems_with_gpu_amd._lib import call, dptr, stream_ptr
torch.cuda.set_device(0)
D = 48620 * 48620
x = torch.empty(D, dtype=torch.float64, device="cuda")
Lanczos(48620, 48620).random_vector(x)
out = []
for one_stream in (False, True):
    if one_stream:
        os.environ["ED_EE_PARALLEL_MIN"] = str(1 << 40)
    s = np.zeros(1)
    call("ed_entanglement_entropy_vn", 18, 9, 9, (1 << 9) - 1, 0, dptr(x), s.ctypes.data_as(C.POINTER(C.c_double)),
         stream_ptr())
    out.append(s[0])
print("S", repr(float(out[0])), repr(float(out[1])), flush=True)
"""


def test_first_call_of_process_18_sites(tmp_path):
    """The first multi-worker entanglement call of a process (the one that warms rocSOLVER up on the
    calling thread before the worker threads start) at 18 sites, 9|9: the large blocks are solved
    before the warm-up, whose values must not land on theirs.  Fresh process; the same S_vN as the
    one-stream path."""
    import os
    import subprocess
    import sys
    script = tmp_path / "first_call.py"
    script.write_text(_FIRST_CALL_SCRIPT)
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    print("entropy 18 sites, first call of a fresh process", flush=True)
    r = subprocess.run([sys.executable, str(script), repo], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    line = [l for l in r.stdout.splitlines() if l.startswith("S ")][-1]
    s_workers, s_one = (float(v) for v in line.split()[1:])
    assert 0.0 < s_one < 18 * math.log(4.0)
    assert abs(s_workers - s_one) <= 1e-10, (s_workers, s_one)

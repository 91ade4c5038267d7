#!/usr/bin/env python3
"""H.v throughput benchmark: 18-site Kagome Hubbard, N_up = N_dn = 9, FP64.

One step = one H.v (Siro-Harju split) over the full sector, on the slab layout the Lanczos
solver keeps its vectors in (bit-identical to the reference layout's H.v, which is timed and
reported beside it),
D = 48,620^2 = 2,363,904,400 amplitudes, x = the normalized seed-235 start vector,
bonds of the 2x3-cell Kagome cluster (KagomeLattice18y, drude_v2_cal_Kagome_randomHop.cu:712-730)
with matrix element -1 and U = 8.  Inputs are resident in HBM before timing starts.

value = algorithmic bytes (16 D per H.v: x read once, y written once) / mean step time.
The JSON line also carries the roofline of the H.v kernel (HBM bound, MI355X peak
8 TB/s) and a CPU baseline: the oracle's H.v (a C restatement of the reference
kernel, OpenMP over rows) timed on a bounded sample of rows of the same workload.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "H·v HBM GB/s vs peak (18-site Kagome); batched Drude weights/s vs plain @1e-8"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(REPO, "profiles", "hv_slab_kagome18y_traffic.txt")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def read_traffic(shape):
    """HBM bytes per H.v launch for the slab tile shape (tile_rows, gathers_in_flight, tile_order)
    from the committed rocprofv3 PMC summary (None when that shape was not measured)."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    vals = {}
    for line in open(TRAFFIC_FILE):
        if "=" in line and not line.lstrip().startswith("#"):
            k, v = line.split("=", 1)
            vals[k.strip().replace(" ", "")] = v.strip()
    key = "hbm_bytes_per_launch[%d,%d,%d]" % tuple(shape)
    try:
        return float(vals[key])
    except (KeyError, ValueError):
        return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(H_or, x_host, dim_dn, budget_s=15.0):
    """Oracle H.v on a bounded row sample; returns (GB/s by the same 16 B/element count, rows, secs)."""
    import numpy as np
    from oracle import oracle as O
    y = np.empty_like(x_host)  # only the sampled rows are touched
    rows = 16
    t0 = time.perf_counter()
    H_or.apply_rows(x_host, y, 0, rows)
    dt = time.perf_counter() - t0
    per_row = max(dt / rows, 1e-6)
    rows2 = int(min(H_or.dim_up - rows, max(16, budget_s / per_row)))
    t0 = time.perf_counter()
    H_or.apply_rows(x_host, y, rows, rows2)
    dt2 = time.perf_counter() - t0
    elems = rows2 * dim_dn
    return 16.0 * elems / dt2 / 1e9, rows2, dt2, O.num_threads()


def ground_state_line(H, hq, x, y, N, Q, reorth=False, slab_shape=None):
    """Lanczos ground state (reference convergence test) of the benchmark Hamiltonian from the
    seed-235 start, Ritz vector rebuilt by recomputation, then S_vN of sites 0..N/2-1.  reorth:
    a second Gram-Schmidt pass of every f against v_N (DGKS), the reorthogonalisation that three
    resident vectors allow.  slab_shape: the slab H.v launch shape (tile rows, gathers in flight,
    tile order) the H.v autotuner chose."""
    import ctypes as C
    import numpy as np
    import torch
    from exact_diagonalization_for_fermionic_systems_with_gpu_amd import Lanczos
    from exact_diagonalization_for_fermionic_systems_with_gpu_amd._lib import call, dptr, stream_ptr
    lz = Lanczos(hq.count, hq.count, reorth=reorth)
    lz.random_vector(x)
    H.slab_layout(True)  # the recurrence runs on the slab layout (start and Ritz vectors converted)
    if slab_shape is not None:
        H.slab_set_shape(*slab_shape)
    torch.cuda.synchronize()
    cur, peak = C.c_int64(), C.c_int64()
    call("ed_memory_stats", C.byref(cur), C.byref(peak), 1)
    t0 = time.perf_counter()
    e0 = lz.eigen(H, x, y)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # the omega monitor's estimate of the largest |v_j . v_k| over the run (Simon's recurrence on
    # the run's a and b): semi-orthogonality is <= sqrt(eps) = 1.5e-8
    lz_omega = lz.so_stats()["omega_max"]
    call("ed_memory_stats", C.byref(cur), C.byref(peak), 0)
    D = hq.count * hq.count
    # one forward step: the fixed-length recurrence (ed_lanczos_tridiagonal, same step kernels
    # without the convergence checks) at 2 and 12 steps; the difference is 10 steps
    f0 = torch.empty_like(x)
    lz.random_vector(f0)
    a_out, b_out = np.zeros(16), np.zeros(16)
    tt = []
    for ns in (2, 12):
        torch.cuda.synchronize()
        t_a = time.perf_counter()
        call("ed_lanczos_tridiagonal", H.handle, 0, ns, dptr(f0), a_out.ctypes.data_as(C.POINTER(C.c_double)),
             b_out.ctypes.data_as(C.POINTER(C.c_double)), stream_ptr())
        torch.cuda.synchronize()
        tt.append(time.perf_counter() - t_a)
    ms_step = (tt[1] - tt[0]) / 10 * 1e3
    del f0
    # the caller's start and Ritz vectors plus everything the library held at its peak (the slab
    # layout's tables and scratch, the solver's work vectors); rocBLAS workspaces not included
    peak_bytes = 2 * 8 * D + peak.value
    H.slab_layout(False)
    torch.cuda.empty_cache()
    log(f"[bench] ground state E0 {e0:.15f} in {lz.nsteps} steps, {t1 - t0:.1f} s")
    s = np.zeros(1)
    call("ed_entanglement_stats", None, None, None, 1)
    torch.cuda.synchronize()
    t_e = time.perf_counter()
    call("ed_entanglement_entropy_vn", N, Q, Q, (1 << (N // 2)) - 1, 0, dptr(y), s.ctypes.data_as(C.POINTER(C.c_double)),
         stream_ptr())
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    log(f"[bench] S_vN({N // 2}|{N - N // 2}) {s[0]:.15f} in {t2 - t_e:.1f} s")
    call("ed_entanglement_stats", None, None, None, 1)
    # the same entropy with every block on one stream: each Gram kernel then runs alone, so its
    # event-timed duration is its own (on the worker streams the Gram overlaps other blocks'
    # eigenvalue kernels and its events also count their time)
    s1 = np.zeros(1)
    prev = os.environ.get("ED_EE_PARALLEL_MIN")
    os.environ["ED_EE_PARALLEL_MIN"] = str(1 << 40)
    try:
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        call("ed_entanglement_entropy_vn", N, Q, Q, (1 << (N // 2)) - 1, 0, dptr(y),
             s1.ctypes.data_as(C.POINTER(C.c_double)), stream_ptr())
        torch.cuda.synchronize()
        t4 = time.perf_counter()
    finally:
        if prev is None:
            del os.environ["ED_EE_PARALLEL_MIN"]
        else:
            os.environ["ED_EE_PARALLEL_MIN"] = prev
    log(f"[bench] S_vN on one stream {s1[0]:.15f} in {t4 - t3:.1f} s")
    out = {"E0": e0, "lanczos_steps": lz.nsteps, "lanczos_s": round(t1 - t0, 2), "reorth": bool(reorth),
           "ms_per_forward_step": round(ms_step, 2), "peak_device_gb": round(peak_bytes / 1e9, 2),
           "peak_device_vectors": round(peak_bytes / (8.0 * D), 2),
           "S_vN": float(s[0]), "entropy_s": round(t2 - t_e, 2), "entropy_one_stream_s": round(t4 - t3, 2),
           "S_vN_one_stream_minus_workers": float(s1[0] - s[0]),
           "omega_max_estimate": lz_omega}
    out.update(rho_a_roofline())
    return out


FP64_MFMA_PEAK_TFLOPS = 78.6


def rho_a_roofline():
    """MFMA roofline of the rho_A Gram kernels of the entanglement calls since the last reset
    (library statistics: SYRK FLOPs over the kernels' event-timed durations) and the eigenvalue
    stage's kernel time.  Meaningful when the blocks ran on one stream."""
    import ctypes as C
    from exact_diagonalization_for_fermionic_systems_with_gpu_amd._lib import call
    g_ms, g_fl, e_ms = C.c_double(), C.c_double(), C.c_double()
    call("ed_entanglement_stats", C.byref(g_ms), C.byref(g_fl), C.byref(e_ms), 1)
    tf = g_fl.value / (g_ms.value * 1e-3) / 1e12 if g_ms.value > 0 else 0.0
    log(f"[bench] rho_A Gram: {g_fl.value:.4e} FLOPs in {g_ms.value:.1f} ms of kernels = {tf:.1f} TFLOP/s "
        f"({tf / FP64_MFMA_PEAK_TFLOPS:.3f} of FP64 MFMA peak); eigenvalue stage {e_ms.value / 1e3:.2f} s of kernels")
    return {"rho_a_roofline": {"bound": "mfma", "achieved": round(tf, 2), "peak": FP64_MFMA_PEAK_TFLOPS,
                               "unit": "TFLOP/s", "frac": round(tf / FP64_MFMA_PEAK_TFLOPS, 4),
                               "gram_flops": g_fl.value, "gram_ms": round(g_ms.value, 2)},
            "eigenvalue_stage_s": round(e_ms.value / 1e3, 2)}


def kagome12_line(with_cpu=True):
    """The plain method at the reference's CPU-runnable size: 12-site Kagome, 6/6, U = 4, hopping -1,
    Lanczos ground state (reference convergence test, Ritz vector rebuilt) from the seed-235 start
    on the GPU, and the same solve by the oracle (C restatement of Lanczos<T>::eigen, OpenMP) on the
    host cores, E0 compared to 1e-10."""
    import numpy as np
    import torch
    from exact_diagonalization_for_fermionic_systems_with_gpu_amd import H_TBq, Hubbard, Lanczos
    from exact_diagonalization_for_fermionic_systems_with_gpu_amd import lattices as L
    ops = L.hubbard_ops(L.KAGOME12, t=-1.0)
    hq = H_TBq(12, 6)
    hq.add_ops(ops)
    H = Hubbard(hq, hq, 4.0)
    f = torch.empty(H.dim, dtype=torch.float64, device="cuda")
    ev = torch.empty_like(f)
    lz = Lanczos(hq.count, hq.count)
    for _ in range(2):  # the first solve also loads the kernels
        lz.random_vector(f)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0 = lz.eigen(H, f, ev)
        torch.cuda.synchronize()
        gpu_s = time.perf_counter() - t0
    out = {"E0": e0, "lanczos_steps": lz.nsteps, "gpu_s": round(gpu_s, 4),
           "workload": "kagome12_nup6_ndn6_U4_lanczos_ritz"}
    if with_cpu:
        from oracle import oracle as O
        osp = O.Species(12, 6, ops)
        Ho = O.Hubbard(osp, osp, 4.0)
        t0 = time.perf_counter()
        st, e0o, *_ = Ho.lanczos_eigen(O.random_vector(Ho.dim))
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(cpu_s, 3), "unit": "s per ground state (E0 + Ritz vector)",
                               "cores": O.num_threads(), "kind": "port", "cpu_model": cpu_model(),
                               "sample": "the whole solve", "E0": e0o}
        out["rel_err_vs_oracle"] = abs(e0 - e0o) / abs(e0o)
    log(f"[bench] kagome12 ground state E0 {e0:.12f} in {lz.nsteps} steps, {gpu_s:.3f} s on the GPU"
        + (f", {out['cpu_baseline']['value']} s on {out['cpu_baseline']['cores']} host cores" if with_cpu else ""))
    return out


def entropy_only_line(N=18, Q=9, seed=5):
    """S_vN(N/2 | N/2) of a seeded random normalized state of the N-site, Q/Q sector: the rho_A
    blocks (FP64 MFMA Gram) and their eigenvalues alone, to profile them without a ground state,
    every block on one stream (no worker threads).  The block sizes, and so the Gram FLOPs
    (2.789e13 at 18 sites), do not depend on the state."""
    import ctypes as C
    import numpy as np
    import torch
    from exact_diagonalization_for_fermionic_systems_with_gpu_amd import H_TBq
    from exact_diagonalization_for_fermionic_systems_with_gpu_amd._lib import call, dptr, stream_ptr
    torch.cuda.set_device(0)
    hq = H_TBq(N, Q)
    D = hq.count * hq.count
    g = torch.Generator(device="cuda").manual_seed(seed)
    psi = torch.randn(D, dtype=torch.float64, device="cuda", generator=g)
    psi /= torch.linalg.norm(psi)
    s = np.zeros(1)
    torch.cuda.synchronize()
    call("ed_entanglement_stats", None, None, None, 1)
    os.environ["ED_EE_PARALLEL_MIN"] = str(1 << 40)  # one stream: the Gram kernels run alone
    t0 = time.perf_counter()
    call("ed_entanglement_entropy_vn", N, Q, Q, (1 << (N // 2)) - 1, 0, dptr(psi), s.ctypes.data_as(C.POINTER(C.c_double)),
         stream_ptr())
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    log(f"[bench] S_vN({N // 2}|{N - N // 2}) of a random state {s[0]:.15f} in {t1 - t0:.1f} s")
    out = {"S_vN": float(s[0]), "entropy_s": round(t1 - t0, 2), "dim": D,
           "workload": f"random_state_N{N}_nup{Q}_ndn{Q}_cut{N // 2}"}
    out.update(rho_a_roofline())
    return out


def batched_drude_line(with_cpu=False, use_graph=1):
    """Lieb-12 twist sweep: 13 sectors x 16 fluxes = 208 complex128 Lanczos instances, one Drude
    weight per sector; batched (ed_batch_lowest: the device convergence test every ten steps,
    fifty-step hipGraph replays, or with use_graph 2 every step after the first ten in one graph
    launch) vs the plain per-instance loop on the same GPU, both checked against the oracle's
    Drude weights (tests/golden fixture)."""
    import numpy as np
    from exact_diagonalization_for_fermionic_systems_with_gpu_amd import batched as BT
    gold = json.load(open(os.path.join(REPO, "tests", "golden", "lieb12_drude_oracle.json")))
    Dg = np.array(gold["D"])
    # panel shape of the batch kernels: autotuned once on the sweep's 208 instances (untimed)
    tune = BT.BatchedLanczos([h for (_, _, h) in BT.lieb12_instances(4.0, 16)], True).autotune(steps=30)
    shape = (tune["dn_rows"], tune["dn_panels"], tune["up_rows"], tune["up_panels"])
    log(f"[bench] batch autotune: {tune}")
    BT.lieb12_drude_sweep(batched=True, shape=shape, use_graph=use_graph)  # warm-up (allocations, graph paths)
    BT.lieb12_drude_sweep(batched=False)
    # three timed sweeps of each, alternating; the fastest of each is reported
    runs_b, runs_p = [], []
    for _ in range(3):
        runs_b.append(BT.lieb12_drude_sweep(batched=True, shape=shape, use_graph=use_graph))
        runs_p.append(BT.lieb12_drude_sweep(batched=False))
    E0b, Db, tb, _ = min(runs_b, key=lambda r: r[2])
    E0p, Dp, tp, _ = min(runs_p, key=lambda r: r[2])
    nz = np.abs(Dg) > 1e-6
    # every one of the 208 instances: batched vs plain E0, and E0 even in the flux
    e0_diff = float(max(np.max(np.abs(rb[0] - rp[0])) for rb in runs_b for rp in runs_p))
    e0_asym = float(max(np.max(np.abs(r[0][:, 1:] - r[0][:, :0:-1])) for r in runs_b + runs_p))
    err = float(max(max(np.max(np.abs(r[1][nz] - Dg[nz]) / np.abs(Dg[nz])) for r in runs_b + runs_p), 0.0))
    log(f"[bench] batched Drude sweep {tb:.3f} s (runs {[round(r[2], 3) for r in runs_b]}), plain {tp:.3f} s "
        f"(runs {[round(r[2], 3) for r in runs_p]}), max rel err vs oracle {err:.2e}")
    cpu = None
    if with_cpu:
        # the oracle's complex Lanczos E0 (the restatement of Lanczos<T>::eigen) of the 13 sectors at
        # one flux, Phi_1: 13 of the 208 instances, scaled to the sweep
        import math
        from oracle import oracle as O
        from exact_diagonalization_for_fermionic_systems_with_gpu_amd import lattices as L
        ops = L.lieb_twisted_ops(2 * math.pi / 16)
        t0 = time.perf_counter()
        for nu in range(13):
            Ho = O.Hubbard(O.Species(12, nu, ops, True), O.Species(12, 12 - nu, ops, True), 4.0)
            Ho.lanczos_eigen(O.random_vector(Ho.dim, complex_=True))
        cs = time.perf_counter() - t0
        cpu = {"value": round(13 / (cs * 16), 4), "unit": "Drude weights/s", "cores": O.num_threads(),
               "kind": "port", "cpu_model": cpu_model(),
               "sample": f"oracle E0 of the 13 sectors at one flux ({cs:.1f} s), times 16 fluxes"}
        log(f"[bench] Drude sweep CPU sample: 13 instances in {cs:.1f} s")
    return {"value": round(13 / tb, 3), "unit": "Drude weights/s", "plain": round(13 / tp, 3),
            "cpu_baseline": cpu,
            "speedup": round(tp / tb, 3), "max_rel_err_vs_oracle": err, "instances": 208,
            "max_abs_e0_batched_minus_plain": e0_diff, "max_abs_e0_flux_asymmetry": e0_asym,
            "batch_shape": {k: tune[k] for k in ("dn_rows", "dn_panels", "up_rows", "up_panels")},
            "graph": "fifty-step graph replays" if use_graph == 1 else "one graph launch after the first ten steps",
            "workload": "lieb12_N12_13sectors_16fluxes_U4_complex128"}


def ensemble_line(n_real, ncv=24, workers=1, method="batched"):
    """14-site open chain, 7/7, U = 4, W = 8.  Method "batched": the 4 lowest eigenvalues of all
    realisations from one batched Lanczos launch (each realisation stops once its 4 lowest pass
    the test, at most 600 steps) and their ground states from a batched rebuild
    (ed_batch_lowest_k), a realisation whose values miss the test solved alone as with "lanczos"; "lanczos": E0 of all in one batched launch, then per realisation one Lanczos
    run with selective orthogonalisation, ground state refined by a plain run; "restart": the
    thick-restart solver with ncv = 24 basis vectors.  Then dE0/d eps_i per realisation (`workers`
    at a time on separate streams) and the 7|7 entropy S_vN of every ground state, each block size
    of rho_A solved for all realisations in one batched call."""
    from exact_diagonalization_for_fermionic_systems_with_gpu_amd import disorder as DS
    res = DS.run_ensemble(n_real=n_real, L_sites=14, n_up=7, n_dn=7, W=8.0, nev=4, ncv=ncv, tol=1e-10,
                          log=lambda m: log("[bench] " + m), with_sref=False, workers=workers, method=method)
    log(f"[bench] disorder ensemble ({method}, {workers} worker(s)): {n_real} realisations in {res.seconds:.2f} s "
        f"(batched launch: {res.seconds_batched_e0:.2f} s, {res.fallbacks} solved alone)")
    return {"value": round(res.realisations_per_s, 3), "unit": "realisations/s", "realisations": n_real,
            "batched_s": round(res.seconds_batched_e0, 3), "solved_alone": res.fallbacks,
            "max_abs_e0_batched_minus_eigs": float(abs(res.e0_batched - res.evals[:, 0]).max()),
            "method": method, "workers": workers, "ncv": ncv if method == "restart" else None,
            "workload": "chain14_open_nup7_ndn7_U4_W8_nev4"}


def krylov_line(max_steps=1000, dt=0.02):
    """Krylov<C> quench step (HubMatrix_GPU.cuh:619-728) on the 14-site open chain, 7/7, U = 4,
    W = 8 (realisation 0 of the ensemble): from the product state up = sites 0-6, dn = sites
    7-13, the Lanczos recurrence to max_steps (b stays above 1e-10 here), all eigenpairs of T,
    and the propagation exp(-i H dt) phi0 by the recomputed recurrence; complex128."""
    import numpy as np
    import torch
    from exact_diagonalization_for_fermionic_systems_with_gpu_amd import H_TBq, Krylov, tridiag_eigen
    from exact_diagonalization_for_fermionic_systems_with_gpu_amd import disorder as DS
    H = DS.build_realisation(DS.realisation_eps(0, 14, 8.0))
    hq = H_TBq(14, 7)
    i0 = int(hq.address[0x007f].item()) * H.dimDN + int(hq.address[0x3f80].item())
    phi0 = torch.zeros(H.dim, dtype=torch.complex128, device="cuda")
    phi0[i0] = 1.0
    phi = torch.zeros_like(phi0)
    kr = Krylov(H.dimUP, H.dimDN, max_steps=max_steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kr.tridiagonal(H, phi0)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kr.m, kr.z = tridiag_eigen(kr.a, kr.b)
    t2 = time.perf_counter()
    kr.Krylov_time_evol(H, phi0, phi, dt)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    n = len(kr.a)
    nrm = float(torch.linalg.vector_norm(phi).item())
    log(f"[bench] Krylov quench step: {n} Lanczos steps {t1 - t0:.2f} s, eigenpairs of T {t2 - t1:.2f} s, "
        f"propagation {t3 - t2:.2f} s, |phi| - 1 = {nrm - 1.0:.1e}")
    return {"value": round(1.0 / (t3 - t0), 4), "unit": "Krylov time steps/s", "krylov_dim": n, "dt": dt,
            "tridiagonal_s": round(t1 - t0, 3), "eigenpairs_s": round(t2 - t1, 3), "propagation_s": round(t3 - t2, 3),
            "norm_error": abs(nrm - 1.0), "workload": "chain14_open_nup7_ndn7_U4_W8_krylov1000_complex128"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sites", type=int, default=18, choices=[12, 18])
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched Drude sweep and the disorder-ensemble lines")
    ap.add_argument("--ensemble", type=int, default=128, help="realisations of the 14-site chain ensemble")
    ap.add_argument("--ground-state", action="store_true",
                    help="run the ground-state line even with --no-extras")
    ap.add_argument("--no-ground-state", action="store_true",
                    help="skip the Lanczos ground state of the same workload (Ritz vector rebuilt), its "
                         "half-cluster entropy S_vN (particle-number blocks of rho_A, FP64 MFMA Gram) and peak "
                         "device memory")
    ap.add_argument("--reorth", action="store_true",
                    help="with --ground-state: a DGKS pass of each new Lanczos vector against v_N")
    ap.add_argument("--entropy-only", action="store_true",
                    help="only S_vN(9|9) of a seeded random 18-site 9/9 state (rho_A Gram + block eigenvalues)")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="run only the disorder-ensemble line (--ensemble realisations, --workers threads)")
    ap.add_argument("--workers", type=int, default=1, help="host threads of the disorder-ensemble line")
    ap.add_argument("--ensemble-method", default="batched", choices=["batched", "lanczos", "restart"],
                    help="eigensolver of the disorder-ensemble line")
    ap.add_argument("--drude-graph", type=int, default=1, choices=[1, 2],
                    help="batched Drude sweep: 1 fifty-step graph replays, 2 one graph for the whole solve")
    ap.add_argument("--drude-only", action="store_true",
                    help="run only the batched Drude sweep line (for profiling its kernels)")
    ap.add_argument("--copy-ref", action="store_true",
                    help="also run one device copy of a vector (16 D bytes) as the byte-count reference for "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the H.v path here is single-GPU (multi-GPU sharding is a design "
              "note only), so --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    if args.drude_only:
        print(json.dumps({"batched_drude": batched_drude_line(use_graph=args.drude_graph)}), flush=True)
        return
    if args.ensemble_only:
        print(json.dumps({"disorder_ensemble": ensemble_line(args.ensemble, workers=args.workers,
                                                             method=args.ensemble_method)}), flush=True)
        return
    if args.entropy_only:
        print(json.dumps({"entropy": entropy_only_line()}), flush=True)
        return

    import torch
    from exact_diagonalization_for_fermionic_systems_with_gpu_amd import H_TBq, Hubbard, Lanczos
    from exact_diagonalization_for_fermionic_systems_with_gpu_amd import lattices as L

    torch.cuda.set_device(0)
    cl = L.KAGOME18Y if args.sites == 18 else L.KAGOME12
    N = cl.n_sites
    Q = N // 2
    U = 8.0 if args.sites == 18 else 4.0
    ops = L.hubbard_ops(cl, t=-1.0)
    log(f"[bench] building basis/tables N={N} Q={Q}")
    hq = H_TBq(N, Q)
    hq.add_ops(ops)
    H = Hubbard(hq, hq, U)
    D = H.dim
    log(f"[bench] D = {D} ({D * 8 / 1e9:.2f} GB per vector)")
    x = torch.empty(D, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    Lanczos(hq.count, hq.count).random_vector(x)
    torch.cuda.synchronize()
    tuning = H.autotune(x, y, reps=2)
    log(f"[bench] autotuned tile shapes: {tuning}")

    for _ in range(args.warmup):
        H.Hubbard_mv(x, y)
    torch.cuda.synchronize()
    log(f"[bench] warmup done ({args.warmup} steps)")

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record()
        H.Hubbard_mv(x, y)
        ends[k].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = sorted(s.elapsed_time(e) for s, e in zip(starts, ends))  # ms
    rm_mean_ms = sum(per) / len(per)
    rm_median_ms = per[len(per) // 2]
    alg_bytes = 16.0 * D
    log(f"[bench] H.v on the reference layout: mean {rm_mean_ms:.3f} ms, median {rm_median_ms:.3f} ms, "
        f"wall/step {wall / args.steps * 1e3:.3f} ms")

    # the solver's layout: the same H.v (bit for bit) on the slab layout that ed_lanczos_eigen
    # keeps its vectors in; the input is converted once before timing, as the solver does
    t0 = time.perf_counter()
    H.slab_layout(True)
    xl = H.to_slab(x)
    yl = torch.empty_like(xl)
    torch.cuda.synchronize()
    slab_setup_s = time.perf_counter() - t0
    slab_tuning = H.slab_autotune(xl, yl, reps=2)
    log(f"[bench] slab autotune: {slab_tuning}")
    for _ in range(max(1, args.warmup)):  # at least one, for the bitwise check below
        H.apply_slab(xl, yl)
    same = bool(torch.equal(H.from_slab(yl), y))
    log(f"[bench] slab layout ready in {slab_setup_s:.2f} s; H.v bitwise equal to the reference layout: {same}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record()
        H.apply_slab(xl, yl)
        ends[k].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = sorted(s.elapsed_time(e) for s, e in zip(starts, ends))  # ms
    mean_ms = sum(per) / len(per)
    median_ms = per[len(per) // 2]
    gbs = alg_bytes / (mean_ms * 1e-3) / 1e9
    log(f"[bench] H.v on the slab layout: mean {mean_ms:.3f} ms, median {median_ms:.3f} ms, "
        f"wall/step {wall / args.steps * 1e3:.3f} ms")
    # gather ceiling: each pass's line gathers alone (same tiles, tile order and entry tables, no
    # arithmetic, no streams); pass 1 reads the transposed copy the last H.v left in the scratch
    ceil_ms = []
    for p in (0, 1):
        H.slab_gather_ceiling(xl, p)
        torch.cuda.synchronize()
        for k in range(args.steps):
            starts[k].record()
            H.slab_gather_ceiling(xl, p)
            ends[k].record()
        torch.cuda.synchronize()
        ceil_ms.append(sum(s.elapsed_time(e) for s, e in zip(starts, ends)) / args.steps)
    log(f"[bench] gather ceiling: UP {ceil_ms[0]:.3f} ms, DN {ceil_ms[1]:.3f} ms; "
        f"H.v / ceiling = {mean_ms / sum(ceil_ms):.3f}")
    conv = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    conv[0].record()
    H.to_slab(x, xl)
    conv[1].record()
    torch.cuda.synchronize()
    conv_ms = conv[0].elapsed_time(conv[1])
    del xl, yl
    H.slab_layout(False)
    torch.cuda.empty_cache()
    if not same:
        raise SystemExit("slab-layout H.v differs from the reference-layout H.v")

    if args.copy_ref:
        torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        log(f"[bench] copy reference: {16 * D} bytes moved by one copy kernel")
    traffic = read_traffic((slab_tuning["tile_rows"], slab_tuning["gathers_in_flight"], slab_tuning["tile_order"]))
    out = {
        "metric": METRIC,
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seed-235 minstd_rand start vector, the reference's GenRand)",
        "config": {"workload": f"hubbard_hv_{cl.name}_nup{Q}_ndn{Q}_U{U:g}", "sites": N, "dim": D,
                   "layout": "slab (the Lanczos solver's vector layout; bit-identical to the reference layout)",
                   "hv_median_ms": round(median_ms, 4), "hv_mean_ms": round(mean_ms, 4),
                   "hv_reference_layout_mean_ms": round(rm_mean_ms, 4),
                   "hv_reference_layout_gbs": round(alg_bytes / (rm_mean_ms * 1e-3) / 1e9, 2),
                   "reference_to_slab_conversion_ms": round(conv_ms, 3),
                   "gather_ceiling_ms": {"up": round(ceil_ms[0], 4), "dn": round(ceil_ms[1], 4),
                                         "total": round(sum(ceil_ms), 4)},
                   "hv_over_gather_ceiling": round(mean_ms / sum(ceil_ms), 4),
                   "hv_tiles": {k: v for k, v in tuning.items() if k != "ms_per_apply"},
                   "slab_tiles": {k: v for k, v in slab_tuning.items() if k != "ms_per_apply"}},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        import numpy as np
        from oracle import oracle as O
        log("[bench] cpu baseline: oracle H.v on a row sample")
        osp = O.Species(N, Q, ops)
        Hor = O.Hubbard(osp, osp, U)
        x_host = np.empty(D, dtype=np.float64)
        torch.from_numpy(x_host).copy_(x)
        cg, rows, secs, thr = cpu_baseline(Hor, x_host, hq.count)
        out["cpu_baseline"] = {"value": round(cg, 3), "unit": "GB/s", "cores": thr, "kind": "port",
                               "cpu_model": cpu_model(),
                               "sample": f"oracle H.v (C restatement of Hubbard_mv_), rows {16}..{16 + rows} "
                                         f"of {hq.count} x {hq.count} ({secs:.1f} s), same 16 B/element count"}
        del x_host
    if args.ground_state or not (args.no_extras or args.no_ground_state):
        out["ground_state"] = ground_state_line(
            H, hq, x, y, N, Q, reorth=args.reorth,
            slab_shape=(slab_tuning["tile_rows"], slab_tuning["gathers_in_flight"], slab_tuning["tile_order"]))
    if not args.no_extras:
        del x, y
        torch.cuda.empty_cache()

        def extra(name, fn):
            # a Python-level failure of one secondary line is recorded in its place; the headline
            # line above is printed either way
            try:
                out[name] = fn()
            except Exception as e:  # noqa: BLE001
                log(f"[bench] {name} failed: {type(e).__name__}: {e}")
                out[name] = {"error": f"{type(e).__name__}: {e}"}

        extra("batched_drude", lambda: batched_drude_line(with_cpu=not args.no_cpu, use_graph=args.drude_graph))
        extra("kagome12_ground_state", lambda: kagome12_line(with_cpu=not args.no_cpu))
        extra("disorder_ensemble", lambda: ensemble_line(args.ensemble, workers=args.workers,
                                                         method=args.ensemble_method))
        extra("krylov_quench", krylov_line)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

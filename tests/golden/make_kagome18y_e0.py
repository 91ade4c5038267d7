"""Writes kagome18y_e0_oracle.json: the ground-state energy of the 18-site Kagome Hubbard model
(2x3 cells, KagomeLattice18y bonds of drude_v2_cal_Kagome_randomHop.cu:712-730, matrix element
-1, no on-site term), N_up = N_dn = 9, U = 8, D = 2,363,904,400, from the CPU oracle's
two-vector Lanczos (same recurrence, seed-235 start and convergence test as Lanczos<T>::eigen,
HubMatrix_GPU.cuh:561-613) with tol 1e-12.  Needs ~38 GB of host memory and hours of CPU time:

    OMP_NUM_THREADS=6 python tests/golden/make_kagome18y_e0.py
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from exact_diagonalization_for_fermionic_systems_with_gpu_amd import lattices as L  # noqa: E402


def main():
    cl = L.KAGOME18Y
    ops = L.hubbard_ops(cl, t=-1.0)
    sp = O.Species(18, 9, ops)
    H = O.Hubbard(sp, sp, 8.0)
    t0 = time.time()
    st, e0, a, b, n = H.lanczos_lowest_2vec(maxN=1000, tol=1e-12, seed=235, verbose=True)
    out = {"lattice": cl.name, "n_up": 9, "n_dn": 9, "U": 8.0, "t": -1.0, "dim": H.dim, "tol": 1e-12,
           "status": st, "E0": e0, "nsteps": n, "a": list(a[:40]), "b": list(b[:41]),
           "threads": O.num_threads(), "seconds": round(time.time() - t0, 1)}
    with open(os.path.join(HERE, "kagome18y_e0_oracle.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out["E0"], out["nsteps"], out["seconds"])


if __name__ == "__main__":
    main()

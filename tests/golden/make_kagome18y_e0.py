"""Writes the 18-site Kagome ground-state fixtures from the CPU oracle's two-vector Lanczos (the
recurrence and convergence test of Lanczos<T>::eigen, HubMatrix_GPU.cuh:561-613).

Model: Kagome Hubbard on 2x3 cells (KagomeLattice18y bonds of drude_v2_cal_Kagome_randomHop.cu:
712-730), matrix element -1, no on-site term, N_up = N_dn = 9, U = 8, D = 2,363,904,400, from the
seed-235 start vector.

Two fixtures:
- kagome18y_lanczos100_oracle.json (default): alpha[0..99], beta[0..100] and the lowest Ritz value
  of T_100 of a fixed 100-step recurrence (~33 min on 7 cores).
- kagome18y_e0_oracle.json (--converged): the recurrence run to convergence with the reference's
  test (check every 10 steps, tol = 0.01 sqrt(eps), |b_{N+1} z_N| < tol anorm and
  |z_0| > tol anorm): E0, the step count, alpha and beta (~2 h on 7 cores).

Both need ~38 GB of host memory:

    OMP_NUM_THREADS=7 python tests/golden/make_kagome18y_e0.py [--converged]
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from exact_diagonalization_for_fermionic_systems_with_gpu_amd import lattices as L  # noqa: E402

NSTEPS = 100
MAXN = 1000


def main():
    converged = "--converged" in sys.argv[1:]
    cl = L.KAGOME18Y
    ops = L.hubbard_ops(cl, t=-1.0)
    sp = O.Species(18, 9, ops)
    H = O.Hubbard(sp, sp, 8.0)
    t0 = time.time()
    if converged:
        st, theta, a, b, n = H.lanczos_lowest_2vec(maxN=MAXN, seed=235, verbose=True)
        assert st == 0, st
        name = "kagome18y_e0_oracle.json"
    else:
        st, theta, a, b, n = H.lanczos_lowest_2vec(maxN=NSTEPS, tol=0.0, seed=235, verbose=True)
        assert n == NSTEPS
        name = "kagome18y_lanczos100_oracle.json"
    out = {"lattice": cl.name, "n_up": 9, "n_dn": 9, "U": 8.0, "t": -1.0, "dim": H.dim, "nsteps": n,
           "ritz_lowest": theta, "alpha": list(map(float, a[:n])), "beta": list(map(float, b[:n + 1])),
           "threads": O.num_threads(), "seconds": round(time.time() - t0, 1)}
    if converged:
        out["E0"] = theta
        out["tol"] = 0.01 * float(__import__("numpy").sqrt(__import__("numpy").finfo(float).eps))
    with open(os.path.join(HERE, name), "w") as f:
        json.dump(out, f, indent=1)
    print(theta, n, out["seconds"])


if __name__ == "__main__":
    main()

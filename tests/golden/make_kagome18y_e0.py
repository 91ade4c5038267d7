"""Writes kagome18y_lanczos100_oracle.json: the first 100 Lanczos steps of the 18-site Kagome
Hubbard model (2x3 cells, KagomeLattice18y bonds of drude_v2_cal_Kagome_randomHop.cu:712-730,
matrix element -1, no on-site term), N_up = N_dn = 9, U = 8, D = 2,363,904,400, from the seed-235
start vector, by the CPU oracle's two-vector Lanczos (the recurrence of Lanczos<T>::eigen,
HubMatrix_GPU.cuh:561-613): alpha[0..99], beta[0..100] and the lowest Ritz value of T_100.

The ground state converges slowly at this size (|beta_N z_N| ~ 0.1 after 100 steps), so the
fixture pins a fixed number of steps: any H.v or recurrence error shows up in alpha, beta and
the Ritz value long before convergence.  Needs ~38 GB of host memory and ~1.5 h on 8 cores:

    OMP_NUM_THREADS=8 python tests/golden/make_kagome18y_e0.py
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


def main():
    cl = L.KAGOME18Y
    ops = L.hubbard_ops(cl, t=-1.0)
    sp = O.Species(18, 9, ops)
    H = O.Hubbard(sp, sp, 8.0)
    t0 = time.time()
    st, theta, a, b, n = H.lanczos_lowest_2vec(maxN=NSTEPS, tol=0.0, seed=235, verbose=True)
    assert n == NSTEPS
    out = {"lattice": cl.name, "n_up": 9, "n_dn": 9, "U": 8.0, "t": -1.0, "dim": H.dim, "nsteps": n,
           "ritz_lowest": theta, "alpha": list(map(float, a[:n])), "beta": list(map(float, b[:n + 1])),
           "threads": O.num_threads(), "seconds": round(time.time() - t0, 1)}
    with open(os.path.join(HERE, "kagome18y_lanczos100_oracle.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(theta, n, out["seconds"])


if __name__ == "__main__":
    main()

"""Writes chain14_eigs_oracle.json: for realisations r = 0, 1 of the open 14-site Hubbard chain
(hopping -1, U = 4, on-site eps_i ~ Uniform[-8, 8] from numpy PCG64 seeded 2024 + r, the same
potential for both species, N_up = N_dn = 7, D = 11,778,624) the four lowest eigenvalues from
ARPACK (scipy.sparse.linalg.eigsh, which='SA', tol 1e-13) driven by the CPU oracle's H.v, and
for the ground state: the half-chain (sites 0-6) entropies from the oracle's block spectra and
the site densities <n_i up> + <n_i dn> (= dE0/d eps_i).

    OMP_NUM_THREADS=4 python tests/golden/make_chain14_eigs.py
"""
import json
import os
import sys
import time

import numpy as np
from scipy.sparse.linalg import LinearOperator, eigsh

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from oracle import ee_oracle as EE  # noqa: E402

L, NU, ND, U, T, W, SEED0 = 14, 7, 7, 4.0, 1.0, 8.0, 2024


def ops_for(eps):
    ops = [(0, i, i + 1, -T) for i in range(L - 1)]         # Make_H(i, i+1, -t)
    ops += [(2, i, i, float(eps[i])) for i in range(L)]     # Make_N(i, eps_i)
    return ops


def main():
    out = {"L": L, "n_up": NU, "n_dn": ND, "U": U, "t": T, "W": W, "seed0": SEED0, "realisations": []}
    for r in (0, 1):
        eps = np.random.default_rng(SEED0 + r).uniform(-W, W, L)
        ops = ops_for(eps)
        up, dn = O.Species(L, NU, ops), O.Species(L, ND, ops)
        H = O.Hubbard(up, dn, U)
        t0 = time.time()
        A = LinearOperator((H.dim, H.dim), matvec=lambda x: H.apply(np.ascontiguousarray(x).ravel()),
                           dtype=np.float64)
        v0 = O.random_vector(H.dim, 235)
        w, V = eigsh(A, k=4, which="SA", tol=1e-13, v0=v0, ncv=24)
        order = np.argsort(w)
        w, V = w[order], V[:, order]
        psi = V[:, 0] / np.linalg.norm(V[:, 0])
        lam, sig = EE.block_spectra(L, NU, ND, (1 << (L // 2)) - 1, psi)
        svn = float(EE.entropy_from_lambda(lam))
        sref = float(EE.entropy_from_sigma(sig))
        p2 = (psi.reshape(up.count, dn.count)) ** 2
        nu = np.array([[(int(s) >> i) & 1 for i in range(L)] for s in up.state], dtype=float)
        nd = np.array([[(int(s) >> i) & 1 for i in range(L)] for s in dn.state], dtype=float)
        dens = p2.sum(1) @ nu + p2.sum(0) @ nd
        out["realisations"].append({"r": r, "eps": list(eps), "evals": list(map(float, w)), "S_vN": svn,
                                    "S_ref": sref, "dE0_deps": list(map(float, dens)),
                                    "seconds": round(time.time() - t0, 1)})
        print(r, w, svn, sref, time.time() - t0, flush=True)
    with open(os.path.join(HERE, "chain14_eigs_oracle.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()

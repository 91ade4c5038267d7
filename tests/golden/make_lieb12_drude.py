"""Writes tests/golden/lieb12_drude_oracle.json: the CPU oracle's Lanczos ground-state energies
(Lanczos<T>::eigen restated in oracle/ed_oracle.c, complex128) of the twisted Lieb-12 Hubbard
cluster, U = 4, t = 1, all 13 sectors N_up + N_dn = 12, at the fluxes Phi_0 = 0 and
Phi_1 = 2 pi / 16 (E0 is even in Phi, so Phi_15 gives the same value), and the Drude weights
D_s = [E0(Phi_1) - 2 E0(Phi_0) + E0(Phi_15)] / (h^2 * 12), h = 2 pi / 16.

    python tests/golden/make_lieb12_drude.py      (about a minute on 8 cores)
"""
import json
import math
import os
import sys

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, REPO)

from oracle import oracle as O  # noqa: E402
from exact_diagonalization_for_fermionic_systems_with_gpu_amd import lattices as L  # noqa: E402

NFLUX, U = 16, 4.0


def main():
    E = {}
    for k in (0, 1):
        ops = L.lieb_twisted_ops(2 * math.pi * k / NFLUX)
        for nu in range(13):
            nd = 12 - nu
            H = O.Hubbard(O.Species(12, nu, ops, True), O.Species(12, nd, ops, True), U)
            st, e0, *_ = H.lanczos_eigen(O.random_vector(H.dim, complex_=True))
            E[f"{nu}_{k}"] = e0
            print(nu, k, H.dim, st, e0, flush=True)
    h = 2 * math.pi / NFLUX
    D = [(2 * E[f"{nu}_1"] - 2 * E[f"{nu}_0"]) / (h * h * 12) for nu in range(13)]
    out = {"U": U, "t": 1.0, "nflux": NFLUX, "E0": E, "D": D,
           "formula": "D_s = (E0(Phi_1) - 2 E0(Phi_0) + E0(Phi_15)) / (h^2 * 12), h = 2 pi / 16"}
    json.dump(out, open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "lieb12_drude_oracle.json"), "w"),
              indent=1)


if __name__ == "__main__":
    main()

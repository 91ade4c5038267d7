"""Cluster geometries and the operator lists the reference drivers build on them.

Bond lists are the reference's own (their order matters: it fixes the ELL slot
order and therefore the summation order of H.v), transcribed from the driver
sources cited per lattice.  Each builder returns an op list of
(kind, i, j, value[, phase]) tuples in the order the reference calls
H_TBq::Make_H / Make_J / Make_N.

Kagome "Kinetic"/"Current" projections (d_x^2 and d_x per bond family) follow
drude_v2_cal_Kagome_randomHop.cu:242-302 (K: 1 and 1/4 on x, 3/4 on y; J: 1 and 1/2
on x, sqrt(3)/2 on y).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Tuple

OP_HOP, OP_CURRENT, OP_DENSITY, OP_DENSITY2, OP_HOP_PHASE = range(5)


@dataclass
class Cluster:
    name: str
    n_sites: int
    bonds: List[Tuple[int, int]]
    # bond families for the Kubo weights: x list (i, j, proj_K, proj_J) and y list (i, j)
    x_bonds: List[Tuple[int, int, float, float]] = field(default_factory=list)
    y_bonds: List[Tuple[int, int]] = field(default_factory=list)
    vol: float = 0.0
    onsite_randomhop: bool = True  # Make_N(i, -2 r_i^2) after the bonds


def _zip(a, b):
    assert len(a) == len(b)
    return list(zip(a, b))


def _kagome(name, n, first, second, n_ab, firstY, secondY, vol):
    bonds = _zip(first, second)
    xb = [(i, j, 1.0 if k < n_ab else 0.25, 1.0 if k < n_ab else 0.5) for k, (i, j) in enumerate(bonds)]
    return Cluster(name, n, bonds, xb, _zip(firstY, secondY), vol)


# drude_v2_cal_Kagome_randomHop.cu:794-870 (also ConservedTotalSpin_Kagome_randomHop.cu:140-156)
KAGOME12 = _kagome(
    "kagome12", 12,
    [0, 1, 3, 4, 6, 7, 9, 10, 0, 2, 6, 8, 3, 5, 9, 11, 2, 1, 8, 7, 5, 4, 11, 10],
    [1, 3, 4, 0, 7, 9, 10, 6, 2, 6, 8, 0, 5, 9, 11, 3, 1, 11, 7, 5, 4, 8, 10, 2], 8,
    [0, 2, 6, 8, 3, 5, 9, 11, 1, 11, 7, 5, 4, 8, 10, 2],
    [2, 6, 8, 0, 5, 9, 11, 3, 2, 1, 8, 7, 5, 4, 11, 10],
    8.0 * math.sqrt(3.0))

# drude_v2_cal_Kagome_randomHop.cu:712-792 (2x3 unit cells)
KAGOME18Y = _kagome(
    "kagome18y", 18,
    [0, 1, 3, 4, 6, 7, 9, 10, 12, 13, 15, 16, 0, 2, 6, 8, 12, 14, 3, 5, 9, 11, 15, 17,
     2, 1, 17, 16, 8, 7, 5, 4, 14, 13, 11, 10],
    [1, 3, 4, 0, 7, 9, 10, 6, 13, 15, 16, 12, 2, 6, 8, 12, 14, 0, 5, 9, 11, 15, 17, 3,
     1, 17, 16, 8, 7, 5, 4, 14, 13, 11, 10, 2], 12,
    [0, 2, 6, 8, 12, 14, 3, 5, 9, 11, 15, 17, 1, 17, 16, 8, 7, 5, 4, 14, 13, 11, 10, 2],
    [2, 6, 8, 12, 14, 0, 5, 9, 11, 15, 17, 3, 2, 1, 17, 16, 8, 7, 5, 4, 14, 13, 11, 10],
    12.0 * math.sqrt(3.0))

# drude_v2_cal_Kagome_randomHop.cu:630-710 (3x2 unit cells)
KAGOME18X = _kagome(
    "kagome18x", 18,
    [0, 1, 3, 4, 6, 7, 9, 10, 12, 13, 15, 16, 0, 2, 9, 11, 3, 5, 12, 14, 6, 8, 15, 17,
     2, 1, 14, 13, 8, 7, 11, 10, 5, 4, 17, 16],
    [1, 3, 4, 6, 7, 0, 10, 12, 13, 15, 16, 9, 2, 9, 11, 0, 5, 12, 14, 3, 8, 15, 17, 6,
     1, 14, 13, 8, 7, 11, 10, 5, 4, 17, 16, 2], 12,
    [0, 2, 9, 11, 3, 5, 12, 14, 6, 8, 15, 17, 1, 14, 13, 8, 7, 11, 10, 5, 4, 17, 16, 2],
    [2, 9, 11, 0, 5, 12, 14, 3, 8, 15, 17, 6, 2, 1, 14, 13, 8, 7, 11, 10, 5, 4, 17, 16],
    12.0 * math.sqrt(3.0))

# drude_v2_cal_Kagome_randomHop.cu:550-628 (tilted 18-site cluster, VOL by Heron's formula, :20)
KAGOME18 = _kagome(
    "kagome18", 18,
    [0, 1, 8, 9, 15, 17, 3, 4, 11, 6, 12, 14, 3, 2, 8, 7, 12, 13, 0, 5, 11, 10, 15, 16,
     1, 2, 4, 5, 6, 7, 9, 10, 13, 14, 16, 17],
    [1, 8, 9, 15, 17, 0, 4, 11, 6, 12, 14, 3, 2, 8, 7, 12, 13, 0, 5, 11, 10, 15, 16, 3,
     2, 4, 5, 1, 7, 9, 10, 6, 14, 16, 17, 13], 12,
    [3, 2, 8, 7, 12, 13, 0, 5, 11, 10, 15, 16, 2, 4, 5, 1, 7, 9, 10, 6, 14, 16, 17, 13],
    [2, 8, 7, 12, 13, 0, 5, 11, 10, 15, 16, 3, 1, 2, 4, 5, 6, 7, 9, 10, 13, 14, 16, 17],
    2.0 * math.sqrt((math.sqrt(13) + math.sqrt(7) + 2) * (math.sqrt(13) + math.sqrt(7) - 2)
                    * (math.sqrt(13) - math.sqrt(7) + 2) * (-math.sqrt(13) + math.sqrt(7) + 2)))

# drude_v2_cal_Kagome_randomHop.cu:222-302
KAGOME24 = _kagome(
    "kagome24", 24,
    [0, 1, 6, 8, 15, 16, 21, 23, 3, 4, 9, 11, 12, 13, 18, 20, 3, 2, 6, 7, 0, 5, 9, 10, 15, 14, 18, 19, 12, 17,
     21, 22, 1, 2, 4, 5, 7, 8, 10, 11, 13, 14, 16, 17, 19, 20, 22, 23],
    [1, 6, 8, 15, 16, 21, 23, 0, 4, 9, 11, 12, 13, 18, 20, 3, 2, 6, 7, 12, 5, 9, 10, 15, 14, 18, 19, 0, 17, 21,
     22, 3, 2, 4, 5, 1, 8, 10, 11, 7, 14, 16, 17, 13, 20, 22, 23, 19], 16,
    [3, 2, 6, 7, 0, 5, 9, 10, 15, 14, 18, 19, 12, 17, 21, 22, 2, 4, 5, 1, 8, 10, 11, 7, 14, 16, 17, 13, 20, 22,
     23, 19],
    [2, 6, 7, 12, 5, 9, 10, 15, 14, 18, 19, 0, 17, 21, 22, 3, 1, 2, 4, 5, 7, 8, 10, 11, 13, 14, 16, 17, 19, 20,
     22, 23],
    16.0 * math.sqrt(3.0))

# drude_v2_cal_Kagome_randomHop.cu:304-384
KAGOME24A = _kagome(
    "kagome24a", 24,
    [0, 1, 3, 4, 6, 7, 21, 22, 9, 10, 12, 13, 15, 16, 18, 19, 0, 2, 9, 11, 21, 23, 18, 20, 6, 8, 15, 17, 3, 5,
     12, 14, 2, 1, 17, 16, 23, 22, 14, 13, 8, 7, 11, 10, 5, 4, 20, 19],
    [1, 3, 4, 6, 7, 21, 22, 0, 10, 12, 13, 15, 16, 18, 19, 9, 2, 9, 11, 21, 23, 18, 20, 6, 8, 15, 17, 3, 5, 12,
     14, 0, 1, 17, 16, 23, 22, 14, 13, 8, 7, 11, 10, 5, 4, 20, 19, 2], 16,
    [0, 2, 9, 11, 21, 23, 18, 20, 6, 8, 15, 17, 3, 5, 12, 14, 1, 17, 16, 23, 22, 14, 13, 8, 7, 11, 10, 5, 4, 20,
     19, 2],
    [2, 9, 11, 21, 23, 18, 20, 6, 8, 15, 17, 3, 5, 12, 14, 0, 2, 1, 17, 16, 23, 22, 14, 13, 8, 7, 11, 10, 5, 4,
     20, 19],
    16.0 * math.sqrt(3.0))

# drude_v2_cal_Kagome_randomHop.cu:386-466
KAGOME24X = _kagome(
    "kagome24x", 24,
    [0, 1, 3, 4, 6, 7, 9, 10, 12, 13, 15, 16, 18, 19, 21, 22, 0, 2, 12, 14, 3, 5, 15, 17, 6, 8, 18, 20, 9, 11,
     21, 23, 2, 1, 17, 16, 8, 7, 23, 22, 5, 4, 20, 19, 11, 10, 14, 13],
    [1, 3, 4, 6, 7, 9, 10, 0, 13, 15, 16, 18, 19, 21, 22, 12, 2, 12, 14, 0, 5, 15, 17, 3, 8, 18, 20, 6, 11, 21,
     23, 9, 1, 17, 16, 8, 7, 23, 22, 2, 4, 20, 19, 11, 10, 14, 13, 5], 16,
    [0, 2, 12, 14, 3, 5, 15, 17, 6, 8, 18, 20, 9, 11, 21, 23, 1, 17, 16, 8, 7, 23, 22, 2, 4, 20, 19, 11, 10, 14,
     13, 5],
    [2, 12, 14, 0, 5, 15, 17, 3, 8, 18, 20, 6, 11, 21, 23, 9, 2, 1, 17, 16, 8, 7, 23, 22, 5, 4, 20, 19, 11, 10,
     14, 13],
    16.0 * math.sqrt(3.0))

# drude_v2_cal_Kagome_randomHop.cu:468-548
KAGOME24Y = _kagome(
    "kagome24y", 24,
    [0, 1, 3, 4, 6, 7, 9, 10, 12, 13, 15, 16, 18, 19, 21, 22, 0, 2, 6, 8, 12, 14, 18, 20, 3, 5, 9, 11, 15, 17,
     21, 23, 2, 1, 23, 22, 14, 13, 11, 10, 5, 4, 20, 19, 17, 16, 8, 7],
    [1, 3, 4, 0, 7, 9, 10, 6, 13, 15, 16, 12, 19, 21, 22, 18, 2, 6, 8, 12, 14, 18, 20, 0, 5, 9, 11, 15, 17, 21,
     23, 3, 1, 23, 22, 14, 13, 11, 10, 2, 4, 20, 19, 17, 16, 8, 7, 5], 16,
    [0, 2, 6, 8, 12, 14, 18, 20, 3, 5, 9, 11, 15, 17, 21, 23, 1, 23, 22, 14, 13, 11, 10, 2, 4, 20, 19, 17, 16, 8,
     7, 5],
    [2, 6, 8, 12, 14, 18, 20, 0, 5, 9, 11, 15, 17, 21, 23, 3, 2, 1, 23, 22, 14, 13, 11, 10, 5, 4, 20, 19, 17, 16,
     8, 7],
    16.0 * math.sqrt(3.0))

# TotalSpinExcited_Lieb_dilute.cu:147-159 (2x2 cells x 3 sites; x rings 0-1-3-4 and 6-7-9-10)
LIEB12 = Cluster(
    "lieb12", 12,
    _zip([0, 1, 3, 4, 6, 7, 9, 10, 0, 2, 6, 8, 3, 5, 9, 11],
         [1, 3, 4, 0, 7, 9, 10, 6, 2, 6, 8, 0, 5, 9, 11, 3]),
    onsite_randomhop=False)
LIEB12_X_BONDS = list(range(8))          # the two x rings, 4 bonds each
LIEB12_X_RINGS = [[0, 1, 2, 3], [4, 5, 6, 7]]
LIEB12_PI_X_BONDS = [3, 7]               # sign-flipped in TotalSpinExcited_Lieb_randomHop.cu:184
LIEB12_PI_Y_BONDS = [11, 15]             # sign-flipped in ConservedTotalSpin_Lieb_dilute.cu:138-139

CLUSTERS = {c.name: c for c in (KAGOME12, KAGOME18, KAGOME18X, KAGOME18Y, KAGOME24, KAGOME24A, KAGOME24X,
                                KAGOME24Y, LIEB12)}


def chain(L, periodic=False):
    """Open (or periodic) chain 0-1-...-(L-1): Line10/Line16 of Hubbard_EE_test.cu:102-140."""
    b = [(i, i + 1) for i in range(L - 1)]
    if periodic:
        b.append((L - 1, 0))
    return Cluster(f"chain{L}", L, b, onsite_randomhop=False)


def square_3x4():
    """Squared_3by4 of MBL_test/src/Hubbard_test.cu:126-157 (open 3x4 square)."""
    b = [(0, 1), (1, 2), (2, 3), (4, 5), (5, 6), (6, 7), (8, 9), (9, 10), (10, 11),
         (0, 4), (4, 8), (1, 5), (5, 9), (2, 6), (6, 10), (3, 7), (7, 11)]
    return Cluster("square3x4", 12, b, onsite_randomhop=False)


# ---------------------------------------------------------------------------------------
# op-list builders (each mirrors one driver function)
# ---------------------------------------------------------------------------------------
def hubbard_ops(cl: Cluster, t=-1.0, r=None, onsite=None):
    """KagomeLattice* (drude_v2_cal_Kagome_randomHop.cu:712-730 and siblings):
    Make_H(i, j, -r_i r_j) per bond, then Make_N(i, -2 r_i^2) per site (randomHop).
    With r = None the bond value is t (uniform hopping).  onsite=False drops the
    Make_N terms (the randomOnlyHop drivers, drude_v2_cal_Kagome_randomOnlyHop.cu:352-361)."""
    if onsite is None:
        onsite = cl.onsite_randomhop and r is not None
    ops = []
    for (i, j) in cl.bonds:
        v = t if r is None else -r[i] * r[j]
        ops.append((OP_HOP, i, j, float(v)))
    if onsite:
        for i in range(cl.n_sites):
            ri = 1.0 if r is None else r[i]
            ops.append((OP_DENSITY, i, i, float(-2.0 * ri * ri)))
    return ops


def kinetic_ops(cl: Cluster, r=None):
    """Kagome Kinetic* (drude_v2_cal_Kagome_randomHop.cu:810-839): K_x (d_x^2 weights) and K_y (3/4)."""
    rr = (lambda i: 1.0) if r is None else (lambda i: r[i])
    kx = [(OP_HOP, i, j, float(-rr(i) * rr(j) * pk)) for (i, j, pk, pj) in cl.x_bonds]
    ky = [(OP_HOP, i, j, float(-rr(i) * rr(j) * 0.75)) for (i, j) in cl.y_bonds]
    return kx, ky


def current_ops(cl: Cluster, r=None):
    """Kagome Current* (drude_v2_cal_Kagome_randomHop.cu:841-870): J_x (d_x weights) and J_y (sqrt(3)/2)."""
    rr = (lambda i: 1.0) if r is None else (lambda i: r[i])
    py = math.sqrt(3.0) / 2.0
    jx = [(OP_CURRENT, i, j, float(rr(i) * rr(j) * pj)) for (i, j, pk, pj) in cl.x_bonds]
    jy = [(OP_CURRENT, i, j, float(rr(i) * rr(j) * py)) for (i, j) in cl.y_bonds]
    return jx, jy


def lieb_twisted_ops(phi, t=1.0, rings=LIEB12_X_RINGS, spread=True):
    """Lieb12 with a Peierls twist: total phase phi around every x ring, spread as phi/4 on
    each of its 4 bonds (direction first->second of the bond list); y bonds real.
    Hopping matrix element -t (Make_H(i, j, -1.0) in TotalSpinExcited_Lieb_dilute.cu:155).
    spread=False puts the whole phase on the last bond of each ring (bonds 3 and 7, the ones
    the reference flips), a boundary twist; both are gauge-equivalent, and phi = pi gives the
    reference's sign-flipped x bonds (TotalSpinExcited_Lieb_randomHop.cu:176-188)."""
    xset = {k: len(r) for r in rings for k in r}
    last = {r[-1] for r in rings}
    ops = []
    for k, (i, j) in enumerate(LIEB12.bonds):
        if k in xset:
            # hop i -> j along the ring picks up e^{+i phi/L}: term c+_j c_i
            ph = phi / xset[k] if spread else (phi if k in last else 0.0)
            ops.append((OP_HOP_PHASE, j, i, -t, ph))
        else:
            ops.append((OP_HOP, i, j, -t))
    return ops


def lieb_pi_flipped_ops(t=1.0):
    """LiebLattice12 of TotalSpinExcited_Lieb_randomHop.cu:176-188 with r_i = 1: Make_H(i, j, -t)
    on every bond except x bonds 3 and 7, which get +t (a pi flux through each x ring)."""
    return [(OP_HOP, i, j, t if k in LIEB12_PI_X_BONDS else -t) for k, (i, j) in enumerate(LIEB12.bonds)]

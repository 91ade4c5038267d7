"""CPU checks of the solvers' slab-layout decision (ed_slab_auto_decision, host only): the layout
is built for real sectors of 2^27 amplitudes and more only when its scratch vector and every
layout vector the solver allocates fit with 5 % of the card to spare, so a sector that fits on the
reference layout (two work vectors) never fails on the slab layout for lack of memory."""
from math import comb

from exact_diagonalization_for_fermionic_systems_with_gpu_amd import _lib

GB = 1e9
CARD = 288 * GB


def decide(free, du, dd, nvec, total=CARD):
    return _lib.lib().ed_slab_auto_decision(float(free), float(total), du, dd, nvec)


def slab_len(du, dd):
    return ((dd + 15) // 16) * du * 16


def test_small_sectors_stay_on_reference_layout():
    d = comb(12, 6)
    assert decide(CARD, d, d, 4) == 0          # 853,776 amplitudes
    d = comb(14, 7)
    assert decide(CARD, d, d, 4) == 0          # 1.2e7 amplitudes
    d = comb(16, 8)                             # 1.66e8 >= 2^27 amplitudes, plenty of memory
    assert decide(CARD, d, d, 4) == 1


def test_18_site_ground_state_fits():
    d = comb(18, 9)
    V = 8.0 * slab_len(d, d)
    free = CARD - 2 * 8.0 * d * d               # caller's f and evec resident
    assert decide(free, d, d, 4) == 1
    # the decision counts the scratch copy and all four solver vectors: one vector short fails
    need = 5 * V + 0.05 * CARD
    assert decide(need, d, d, 4) == 1
    assert decide(need - 1.0, d, d, 4) == 0
    assert decide(4 * V + 0.05 * CARD, d, d, 3) == 1


def test_19_site_8_8_falls_back():
    """19 sites, 8/8: V = 45.7 GB.  With f and evec resident, 197 GB are free: enough for the
    reference layout's two work vectors, not for the slab path's five vectors."""
    d = comb(19, 8)
    V = 8.0 * d * d
    free = CARD - 2 * V
    assert free > 2 * V
    assert decide(free, d, d, 4) == 0
    assert decide(free, d, d, 3) == 0 or 4 * 8.0 * slab_len(d, d) + 0.05 * CARD <= free

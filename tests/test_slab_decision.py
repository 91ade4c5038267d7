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
    """16-wide slabs of dn positions, the last one as wide as the remainder (rounded up to even)"""
    r = dd - ((dd + 15) // 16 - 1) * 16
    return ((dd + 15) // 16 - 1) * du * 16 + du * (r + (r & 1))


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


def test_layout_length_has_no_padding_for_even_dims():
    for d in (comb(18, 9), comb(19, 8), comb(12, 6), 70, 48):
        assert slab_len(d, d) == d * d
    assert slab_len(5, comb(8, 0) * 3) == 5 * 4        # 3 dn positions: one padding column
    assert slab_len(7, 33) == 7 * 34


def test_19_site_8_8():
    """19 sites, 8/8: V = 45.7 GB.  With f and evec resident, 197 GB are free.  The slab path with
    its own copies of f and evec (four layout vectors and the scratch) does not fit and the solver
    stays on the reference layout (two work vectors); with the caller's f and evec carrying layout
    vectors (two more plus the scratch: five vectors in all) it fits."""
    d = comb(19, 8)
    V = 8.0 * d * d
    free = CARD - 2 * V
    assert free > 2 * V
    assert decide(free, d, d, 4) == 0
    assert decide(free, d, d, 2) == 1
    assert 5 * V < CARD

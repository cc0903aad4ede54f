"""The row-pair LDS layout of the decoder (two check rows per thread, even Z), restated on
the CPU: for every even lifting size, every thread r < Z/2 and every cyclic shift, the
byte that the device computes from the two-word edge table (col*Z | S << 16, and the
16-bit pair S | (S - T) << 16, added to 2r in both halves by one packed 16-bit add, then
the smaller half) is where the layout stores row (r + shift) mod Z of the column, and
byte ^ 1 holds row (r + Z/2 + shift) mod Z; the layout is a permutation of each
column's bytes."""
import numpy as np
import pytest

import oracle as O

EVEN_Z = [z for z in O.LIFTING_SIZES if z % 2 == 0]


def pair_pos(j, Z):
    """LDS byte of row j within its column (rows j and j + Z/2 adjacent)."""
    h = Z // 2
    return 2 * (j % h) + (1 if j >= h else 0)


def table_words(shift, Z):
    """S and S - T of one edge (column 0), and the host's second row-pair table word."""
    h = Z // 2
    f = 1 if shift >= h else 0
    S = 2 * (shift - f * h) + f
    T = Z - 1 + 2 * f
    return S, S - T, (S & 0xFFFF) | (((S - T) & 0xFFFF) << 16)


@pytest.mark.parametrize("Z", EVEN_Z)
def test_pair_address_matches_layout(Z):
    h = Z // 2
    perm = np.array([pair_pos(j, Z) for j in range(Z)])
    assert sorted(perm.tolist()) == list(range(Z))  # a permutation of the column
    r = np.arange(h, dtype=np.uint64)
    for shift in range(Z):
        S, SmT, g1 = table_words(shift, Z)
        # The 32-bit form: min(2r + S, 2r + S - T) unsigned.
        x = (2 * r + S) & 0xFFFFFFFF
        y = (2 * r + (SmT & 0xFFFFFFFF)) & 0xFFFFFFFF
        assert np.array_equal(np.minimum(x, y), perm[(r + shift) % Z]), (Z, shift)
        # The device's form: 2r in both 16-bit halves plus g1 (lane-wise mod 2^16), then
        # the smaller half.
        lo = (2 * r + (g1 & 0xFFFF)) & 0xFFFF
        hi = (2 * r + (g1 >> 16)) & 0xFFFF
        a = np.minimum(lo, hi)
        assert np.array_equal(a, perm[(r + shift) % Z]), (Z, shift)
        assert np.array_equal(a ^ 1, perm[(r + h + shift) % Z]), (Z, shift)


def test_hard_bit_positions():
    """pair_pos in index form (q + j or q + j - Z + 1) equals the per-row byte."""
    for Z in EVEN_Z:
        for col in (0, 1, 7):
            for j in range(Z):
                q = col * Z + j
                pos = q + j if j < Z // 2 else q + j - Z + 1
                assert pos == col * Z + pair_pos(j, Z)

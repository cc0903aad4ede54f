"""The learned min-sum oracle (oracle.decode_learned) pinned by closed forms: the table
A = 205, B = -1 reproduces round(0.8 m) for every magnitude, so it must decode exactly as
the plain oracle (itself pinned to the reference decoder); A = 0 zeroes every
check-to-variable message, so the soft bits stay the clamped channel values."""
import numpy as np
import pytest

import oracle as O


def _tables(bg, a, b, iters=8):
    M = O.M_rows(bg)
    return np.full((iters, M), a, np.uint16), np.full((iters, M), b, np.int16)


@pytest.mark.parametrize("bg,Z,E,Qm,snr", [(1, 384, 16896, 2, 2.0), (2, 208, 4160, 4, 5.0), (1, 36, 1200, 2, 1.0),
                                          (2, 15, 600, 6, 9.0)])
def test_learned_with_plain_table_equals_plain(bg, Z, E, Qm, snr):
    _, llrs = O.make_codeblocks(12, bg, Z, E, Qm, snr, seed=Z)
    a, b = _tables(bg, 205, -1)
    for i in range(12):
        rm = O.rate_dematch(llrs[i], bg, Z, Qm=Qm)
        r0, b0 = O.decode(rm, bg, Z, max_iter=8)
        r1, b1 = O.decode_learned(rm, bg, Z, a, b, max_iter=8)
        assert r0 == r1 and np.array_equal(b0, b1)


def test_learned_zero_table_keeps_channel_values():
    bg, Z = 2, 52
    K = O.K_b(bg) * Z
    _, llrs = O.make_codeblocks(4, bg, Z, 2400, 2, 3.0, seed=3)
    a, b = _tables(bg, 0, 0)
    for i in range(4):
        rm = O.rate_dematch(llrs[i], bg, Z, Qm=2)
        r, bits = O.decode_learned(rm, bg, Z, a, b, max_iter=3)
        soft = np.concatenate([np.zeros(2 * Z, np.int16), np.clip(rm.astype(np.int16), -64, 64)])[:K]
        hard = np.packbits((soft <= 0).astype(np.uint8))
        assert np.array_equal(bits, hard)
        ok = not (soft == 0).any() and O.crc(O.CRC24B, hard, K) == 0
        assert r == (1 if ok else 0)

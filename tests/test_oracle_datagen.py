"""The generator's CPU statement (oracle.gen_codeblock): message stream, CRC, encoding,
rate matching and channel pinned against known answers, the parity-check equations and
the reference's own encoder + rate matcher (oracle/_ref)."""
import numpy as np
import pytest

import oracle as O

CASES = [(1, 384, 0, O.CRC24B, 2, 0), (2, 208, 16, O.CRC24A, 4, 1), (1, 15, 0, O.CRC16, 6, 2),
         (2, 2, 0, 0, 8, 3), (1, 104, 40, O.CRC24B, 8, 0), (2, 36, 8, O.CRC16, 2, 2)]


def test_message_stream_known_answer():
    # splitmix64 seeded with 0: first output 0xe220a8397b1dcdaf (the published test vector).
    msg, _ = O.gen_codeblock(1, 384, 400, 2, 0.0, seed_msg=0, seed_noise=1, crc_poly=0)
    assert msg[:8].tobytes() == bytes.fromhex("e220a8397b1dcdaf")


@pytest.mark.parametrize("case", CASES)
def test_generated_codeblock_structure(case):
    bg, Z, F, poly, Qm, rv = case
    K, N = O.K_b(bg) * Z, O.N_short(bg) * Z
    E = max(Qm, int(N * 0.9) // Qm * Qm)
    msg, llr = O.gen_codeblock(bg, Z, E, Qm, 0.0, seed_msg=11 + Z, seed_noise=5, rv=rv, F=F, crc_poly=poly)
    bits = np.unpackbits(msg)[:K]
    assert not bits[K - F:].any()
    if poly:
        assert O.crc(poly, msg, K - F) == 0
    cw = O.encode(bits, bg, Z)
    assert O.syndrome_weight(cw, bg, Z) == 0
    # Noise-free: every LLR is +-120 with the sign of the rate-matched bit.
    e = O.rate_match(cw[2 * Z:], E, bg, Z, rv=rv, Qm=Qm, F=F)
    assert np.array_equal(llr, np.where(e == 0, 120, -120).astype(np.int8))
    R = O.ref()
    if R is not None:
        cw_ref = np.zeros(N, np.uint8)
        rm_ref = np.zeros(E, np.uint8)
        R.ref_encode_rate_match(O._p(bits, O._u8p), bg, Z, O._p(cw_ref, O._u8p), O._p(rm_ref, O._u8p), E, rv, Qm, 0,
                                F)
        assert np.array_equal(e, rm_ref)


def test_generated_noise_statistics():
    # Es/N0 = 10 dB QPSK: the quantised LLR magnitude follows 2*sqrt(2)*|x|/n0 clipped to 24;
    # its sign error rate matches Q(sqrt(2 Es/N0 / 2)) = Q(sqrt(10)).
    n0 = 0.1
    errs = tot = 0
    for s in range(8):
        _, llr = O.gen_codeblock(2, 384, 19200, 2, n0, seed_msg=0, seed_noise=100 + s, crc_poly=0)
        _, ref = O.gen_codeblock(2, 384, 19200, 2, 0.0, seed_msg=0, seed_noise=0, crc_poly=0)
        errs += int(((llr <= 0) != (ref <= 0)).sum())
        tot += llr.size
    from math import erfc, sqrt
    p = 0.5 * erfc(sqrt(10.0) / sqrt(2.0))
    assert abs(errs / tot - p) < 4 * sqrt(p / tot) + 1e-5, (errs / tot, p)

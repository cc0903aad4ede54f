"""The oracle's PUSCH soft demodulation and descrambling (oracle/ldpc_oracle.c,
orc_demodulate_soft / orc_scrambling_sequence / orc_pusch_demodulate) pinned to the
reference's own scalar demappers and pseudo-random generator built from source
(oracle/_ref/libref_demod.so), bit-exact, and to what the mathematics fixes: the Gold
sequence's defining recursion, the sign of each LLR on noise-free constellation points
(the mapper's bits), and the special cases (noise variance 0, negative or NaN; symbols
near zero)."""
import numpy as np
import pytest

import oracle as O

MODS = (0, 1, 2, 4, 6, 8)


def _symbols(rng, n):
    iq = (rng.standard_normal(2 * n) * rng.choice([0.05, 0.3, 1, 3], 2 * n)).astype(np.float32)
    nv = rng.choice([1e-4, 1e-2, 0.1, 0.5, 2.0, 0.0, -1.0, np.nan], n).astype(np.float32)
    iq[:64] = 0
    iq[64:128] *= 1e-6
    # Symbols on the interval boundaries of the 64/256-QAM tables (multiples of 2/sqrt(N)).
    for k, N in ((200, 42), (400, 170), (600, 10)):
        m = rng.integers(-9, 10, 100).astype(np.float32)
        iq[k:k + 100] = m * np.float32(2) * (np.float32(1) / np.sqrt(np.float32(N)))
    return iq, nv


def test_scrambling_sequence_is_the_gold_recursion():
    """c(n) = x1(n + 1600) ^ x2(n + 1600) of TS38.211 5.2.1, evaluated with plain lists."""
    for c_init in (0, 1, 0x5a5a5a5, 0x7fffffff):
        N = 1600 + 400
        x1 = [1] + [0] * 30
        x2 = [(c_init >> i) & 1 for i in range(31)]
        for n in range(N):
            x1.append(x1[n + 3] ^ x1[n])
            x2.append(x2[n + 3] ^ x2[n + 2] ^ x2[n + 1] ^ x2[n])
        want = np.array([x1[n + 1600] ^ x2[n + 1600] for n in range(400)], np.uint8)
        assert np.array_equal(O.scrambling_sequence(c_init, 400), want)
        assert np.array_equal(O.scrambling_sequence(c_init, 150, offset=250), want[250:])


def test_scrambling_sequence_matches_reference_build():
    if O.ref_demod() is None:
        pytest.skip("oracle/_ref/libref_demod.so not built (needs the reference tree)")
    for c_init in (0, 1, 17 * 2**15 + 3, 65535 * 2**15 + 1007, 0x7fffffff):
        for off in (0, 1, 31, 1000, 70001):
            a = O.scrambling_sequence(c_init, 4000, off)
            b = O.scrambling_sequence(c_init, 4000, off, use_ref=True)
            assert np.array_equal(a, b), (c_init, off)


@pytest.mark.parametrize("mod", MODS)
def test_demodulate_matches_reference_build(mod):
    if O.ref_demod() is None:
        pytest.skip("oracle/_ref/libref_demod.so not built (needs the reference tree)")
    rng = np.random.default_rng(100 + mod)
    iq, nv = _symbols(rng, 30000)
    a = O.demodulate_soft(iq, nv, mod)
    b = O.demodulate_soft(iq, nv, mod, use_ref=True)
    assert np.array_equal(a, b), (mod, int((a != b).sum()))


@pytest.mark.parametrize("mod", (2, 4, 6, 8))
def test_llr_signs_on_constellation_points(mod):
    """Noise-free symbols of the TS38.211 mapper (oracle modulator), noise variance 0.005 (the
    smallest max-log LLR, 4 / (170 * 0.005), stays above half a quantisation step): every LLR's sign is the
    mapped bit (positive for 0), none is zero."""
    rng = np.random.default_rng(mod)
    nsym = 512
    bits = rng.integers(0, 2, nsym * mod, dtype=np.uint8)
    iq = np.zeros(2 * nsym, np.float32)
    O.lib().orc_modulate(O._p(iq, O.C.POINTER(O.C.c_float)), O._p(bits, O._u8p), bits.size, mod)
    llr = O.demodulate_soft(iq, np.full(nsym, 0.005, np.float32), mod)
    assert (llr != 0).all()
    assert np.array_equal((llr < 0).astype(np.uint8), bits)


def test_special_noise_and_zero_symbols():
    iq = np.array([0.7, -0.2, 0.0, 0.0, 0.3, 0.1], np.float32)
    for mod in (2, 4, 6, 8):
        llr = O.demodulate_soft(iq, np.array([0.0, 1.0, np.nan], np.float32), mod).reshape(3, mod)
        assert (llr[0] == 0).all() and (llr[2] == 0).all()
        if mod > 2:
            assert (llr[1] == 0).all()  # |z|^2 < 1e-9


def test_descrambling_negates_where_sequence_is_one():
    rng = np.random.default_rng(3)
    iq, nv = _symbols(rng, 3000)
    for mod in MODS:
        c_init, off = 4660 * 2**15 + 9, 77
        plain = O.demodulate_soft(iq, nv, mod)
        c = O.scrambling_sequence(c_init, plain.size, off)
        got = O.pusch_demodulate(iq, nv, mod, c_init, off)
        assert np.array_equal(got, np.where(c == 1, -plain, plain).astype(np.int8))

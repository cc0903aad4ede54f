"""The oracle's LLR arithmetic and CRC, pinned by the reference's own KATs
(tests/unittests/phy/upper/log_likelihood_ratio_test.cpp:27-60), by an exhaustive
comparison with the reference build (oracle/_ref) and by closed forms."""
import numpy as np
import pytest

import oracle as O

L = None


def setup_module():
    global L
    L = O.lib()


def add(a, b):
    return int(L.orc_llr_add(a, b))


def psum(a, b):
    return int(L.orc_llr_promotion_sum(a, b))


INF, MAX = 127, 120


def test_reference_kats():
    # log_likelihood_ratio_test.cpp:27-60
    assert add(0, 2) == 2
    assert add(0, -2) == -2
    assert add(2, -2) == 0
    assert psum(0, 2) == 2
    assert add(2, 119) == MAX
    assert add(INF, -2) == INF
    assert add(0, INF) == INF
    assert add(INF, -INF) == 0
    assert add(-100, -100) == -MAX
    assert psum(MAX, MAX) == INF
    assert psum(INF, MAX) == INF


def _rules_add(a, b, promote):
    # special cases in order (log_likelihood_ratio.cpp:22-38), then saturation/promotion
    if a == -b:
        return 0
    if abs(a) > MAX:
        return a
    if abs(b) > MAX:
        return b
    s = a + b
    if abs(s) > MAX:
        return (INF if promote else MAX) * (1 if s > 0 else -1)
    return s


def test_exhaustive_against_rules():
    vals = list(range(-127, 128))
    for a in vals:
        for b in vals:
            assert add(a, b) == _rules_add(a, b, False)
            assert psum(a, b) == _rules_add(a, b, True)


def test_exhaustive_against_reference_build():
    R = O.ref()
    if R is None:
        pytest.skip("reference build (oracle/_ref) not available")
    vals = [v for v in range(-127, 128) if abs(v) <= 120 or abs(v) == 127]
    for a in vals:
        for b in vals:
            assert add(a, b) == R.ref_llr_add(a, b)
            assert psum(a, b) == R.ref_llr_promotion_sum(a, b)


def test_quantize():
    rng = np.random.default_rng(0)
    for v in list(rng.normal(0, 30, 2000)) + [0.0, 20.0, -20.0, 24.0, 100.0, -1e9, 0.0833, -0.0833]:
        for r in (20.0, 24.0):
            c = max(-r, min(r, v))
            x = c / r * 120
            exp = int(np.sign(x) * np.floor(abs(x) + 0.5))
            got = int(L.orc_llr_quantize(v, r))
            assert abs(got - exp) <= 1
            if O.ref() is not None:
                assert got == O.ref().ref_llr_quantize(v, r)


def _poly_crc(bits, poly):
    # closed form: (M(x) * x^order) mod G(x) with Python integers
    order = poly.bit_length() - 1
    m = 0
    for b in bits:
        m = (m << 1) | int(b)
    m <<= order
    while m.bit_length() > order:
        m ^= poly << (m.bit_length() - 1 - order)
    return m


@pytest.mark.parametrize("poly", [O.CRC24A, O.CRC24B, O.CRC16, 0x1b2b117, 0xe21, 0x61])
def test_crc_closed_form(poly):
    rng = np.random.default_rng(poly & 0xff)
    for n in (1, 2, 7, 8, 9, 16, 24, 63, 100, 1000, 8448):
        bits = rng.integers(0, 2, n).astype(np.uint8)
        pk = O.pack_bits(bits)
        r = O.crc(poly, pk, n)
        assert r == _poly_crc(bits, poly)
        # attaching the remainder makes the total remainder zero
        order = poly.bit_length() - 1
        tail = [(r >> (order - 1 - i)) & 1 for i in range(order)]
        full = np.concatenate([bits, np.array(tail, np.uint8)])
        assert O.crc(poly, O.pack_bits(full), full.size) == 0
        if O.ref() is not None:
            assert r == O.ref().ref_crc(poly, O._p(pk, O._u8p), n)

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


def _modulate(fn, bits, Qm):
    import ctypes as C
    iq = np.zeros(2 * (bits.size // Qm), np.float32)
    b = np.ascontiguousarray(bits, np.uint8)
    fn(iq.ctypes.data_as(C.POINTER(C.c_float)), O._p(b, O._u8p), b.size, Qm)
    return iq


def _orc_modulate_fn():
    import ctypes as C
    L = O.lib()
    L.orc_modulate.argtypes = [C.POINTER(C.c_float), O._u8p, C.c_uint, C.c_uint]
    return L.orc_modulate


def _all_patterns(Qm):
    """Every Qm-bit label, bit k of pattern p = (p >> k) & 1, as a (2^Qm, Qm) array."""
    return ((np.arange(1 << Qm)[:, None] >> np.arange(Qm)) & 1).astype(np.uint8)


@pytest.mark.parametrize("Qm", [2, 4, 6, 8])
def test_mapper_matches_reference(Qm):
    """The generator's mapper (orc_modulate, TS38.211 5.1.3-5.1.5) equals the reference's LUT
    modulation mapper (modulation_mapper_lut_impl.cpp, built from source) on every label and
    on a random stream."""
    R = O.ref_demod()
    if R is None:
        pytest.skip("oracle/_ref not built")
    rnd = np.random.default_rng(Qm).integers(0, 2, Qm * 4096).astype(np.uint8)
    for bits in (_all_patterns(Qm).reshape(-1), rnd):
        a = _modulate(_orc_modulate_fn(), bits, Qm)
        b = _modulate(R.ref_modulate, bits, Qm)
        assert np.abs(a - b).max() <= 2e-7, np.abs(a - b).max()


@pytest.mark.parametrize("Qm,esn0_db", [(2, 4.0), (4, 10.0), (6, 16.0), (8, 22.0), (8, 30.0)])
def test_demapper_is_max_log_over_reference_constellation(Qm, esn0_db):
    """The generator's demapper (orc_demap) is the max-log LLR -- (min over the points whose
    bit is 1 of |y - s|^2 minus the same over the points whose bit is 0) / n0 -- over the
    reference mapper's own constellation, quantised by the reference's quantiser with range 24
    (QPSK) or 20 (QAM): brute force over all 2^Qm complex points in float64, against the
    oracle's per-component form in float32 output.  Ties at a quantisation step may round
    either way (|difference| 1, rare)."""
    import ctypes as C
    R, RL = O.ref_demod(), O.ref()
    if R is None or RL is None:
        pytest.skip("oracle/_ref not built")
    labels = _all_patterns(Qm)
    pts = _modulate(R.ref_modulate, labels.reshape(-1), Qm).astype(np.float64).view(np.complex128)
    rng = np.random.default_rng(int(10 * esn0_db) + Qm)
    n0 = float(np.float32(10 ** (-esn0_db / 10)))
    nsym = 3000
    tx = rng.integers(0, 1 << Qm, nsym)
    noise = rng.normal(0, np.sqrt(n0 / 2), (nsym, 2)) @ np.array([1, 1j])
    y = (pts[tx] + noise).astype(np.complex64)
    d = np.abs(y.astype(np.complex128)[:, None] - pts[None, :]) ** 2
    rng_lim = 24.0 if Qm == 2 else 20.0
    expect = np.zeros((nsym, Qm), np.int8)
    for k in range(Qm):
        l = ((d[:, labels[:, k] == 1].min(1) - d[:, labels[:, k] == 0].min(1)) / n0).astype(np.float32)
        expect[:, k] = [RL.ref_llr_quantize(float(v), rng_lim) for v in l]
    L = O.lib()
    L.orc_demap.argtypes = [O._i8p, C.POINTER(C.c_float), C.c_uint, C.c_uint, C.c_float]
    got = np.zeros(nsym * Qm, np.int8)
    yf = np.ascontiguousarray(y.view(np.float32))
    L.orc_demap(O._p(got, O._i8p), yf.ctypes.data_as(C.POINTER(C.c_float)), nsym, Qm, n0)
    diff = np.abs(got.reshape(nsym, Qm).astype(int) - expect.astype(int))
    assert diff.max() <= 1 and (diff != 0).mean() < 1e-3, (diff.max(), (diff != 0).mean())
    # Not a degenerate check: the LLRs cover the quantiser's range.
    assert np.abs(expect).max() >= 60 and len(np.unique(expect)) > 30

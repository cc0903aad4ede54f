"""LDPC encoder + rate matcher on the device (lh_ldpc_encode_rate_match) against the oracle's
encoder and rate matcher, bit-exact: the codeword (N_short*Z bits, one per byte, as
ldpc_encoder_buffer::write_codeblock) and the packed rate-matched bits, over both base
graphs, a sweep of lifting sizes, all redundancy versions, every modulation order, filler
bits and a limited buffer (Nref).  The oracle's encoder and rate matcher are pinned to
the reference's own build in test_oracle_graph_codec.py."""
import numpy as np
import pytest

import ocudu_hip as H
import oracle as O
from gpu_util import require_gpu

pytestmark = pytest.mark.gpu

CASES = []
for bg in (1, 2):
    for Z in (2, 7, 16, 30, 32, 36, 52, 60, 112, 144, 208, 320, 384):
        N = O.N_short(bg) * Z
        K = O.K_b(bg) * Z
        for rv, qm, rate, F, nref in ((0, 2, 0.5, 0, 0), (1, 4, 0.33, 0, 0), (2, 6, 0.8, Z // 2, 0),
                                      (3, 8, 0.6, 0, 0), (0, 1, 0.9, 0, 0), (2, 2, 0.4, 0, (N * 2) // 3)):
            E = max(qm, int(round(K / rate / qm)) * qm)
            if nref and nref < (O.K_b(bg) - 2) * Z + 1:
                continue
            CASES.append((bg, Z, rv, qm, E, F, nref))


def test_encode_rate_match_matches_oracle():
    require_gpu()
    print(f"case encoder + rate matcher, {len(CASES)} codeblocks", flush=True)
    rng = np.random.default_rng(5)
    ctx = H.Context(0, max_codeblocks=len(CASES))
    descs, msgs, moff, ooff, coff = [], [], 0, 0, 0
    exp_out, exp_cw = [], []
    for bg, Z, rv, qm, E, F, nref in CASES:
        K = O.K_b(bg) * Z
        N = O.N_short(bg) * Z
        m = rng.integers(0, 2, K, dtype=np.uint8)
        if F:
            m[K - F:] = 0  # zero fillers at the end of the message
        cw = O.encode(m, bg, Z)[2 * Z:2 * Z + N]
        e = O.rate_match(cw, E, bg, Z, rv=rv, Qm=qm, Nref=nref, F=F)
        exp_cw.append(cw)
        exp_out.append(np.packbits(e))
        pk = np.packbits(m)
        descs.append(H.EncodeDesc(moff, ooff, coff, E, bg, Z, rv, qm, nref, F, 0))
        msgs.append(pk)
        moff += pk.size
        ooff += (E + 7) // 8
        coff += N
    d_msgs = H.DeviceBuffer.from_array(np.concatenate(msgs))
    d_out = H.DeviceBuffer(ooff)
    d_cw = H.DeviceBuffer(coff)
    ctx.encode_rate_match(descs, d_msgs, d_out, d_cw)
    got_out = d_out.download(np.uint8, ooff)
    got_cw = d_cw.download(np.uint8, coff)
    o = c = 0
    for k, (bg, Z, rv, qm, E, F, nref) in enumerate(CASES):
        nb = (E + 7) // 8
        N = O.N_short(bg) * Z
        assert np.array_equal(got_cw[c:c + N], exp_cw[k]), ("codeword", bg, Z, rv, qm, E, F, nref)
        assert np.array_equal(got_out[o:o + nb], exp_out[k]), ("rate matched", bg, Z, rv, qm, E, F, nref)
        o += nb
        c += N
    ctx.close()


def test_encoder_round_trip_through_the_decoder():
    """Encoded and rate-matched on the device, mapped to noise-free LLRs (+-64 for bits 0/1),
    the PUSCH decoder returns the message with its CRC verdict."""
    require_gpu()
    print("case encoder -> decoder round trip", flush=True)
    rng = np.random.default_rng(9)
    bg, Z, E, qm = 1, 384, 9088, 8
    K = 22 * Z
    n = 8
    ctx = H.Context(0, max_codeblocks=n)
    msgs = []
    for _ in range(n):
        m = np.zeros(K, np.uint8)
        kd = K - 24
        m[:kd] = rng.integers(0, 2, kd, dtype=np.uint8)
        r = O.crc(O.CRC24B, np.packbits(m[:kd]), kd)
        for i in range(24):
            m[kd + i] = (r >> (23 - i)) & 1
        msgs.append(np.packbits(m))
    mb = K // 8
    d_msgs = H.DeviceBuffer.from_array(np.concatenate(msgs))
    d_out = H.DeviceBuffer(n * E // 8)
    ctx.encode_rate_match([H.EncodeDesc(i * mb, i * E // 8, 0, E, bg, Z, 0, qm, 0, 0, 0) for i in range(n)],
                          d_msgs, d_out)
    bits = np.unpackbits(d_out.download(np.uint8, n * E // 8))
    llrs = np.where(bits == 0, 64, -64).astype(np.int8)
    d_llr = H.DeviceBuffer.from_array(llrs)
    d_rm = H.DeviceBuffer(n * 66 * Z)
    d_bits = H.DeviceBuffer(n * mb)
    d_res = H.DeviceBuffer(4 * n)
    cbs = [H.PuschCbDesc(H.RmDesc(i * E, i * 66 * Z, E, bg, Z, 0, qm, 0, 0), i * mb, 24, H.LH_CRC24B)
           for i in range(n)]
    ctx.pusch_decode_batch(cbs, d_llr, d_rm, d_bits, d_res, nof_ldpc_iterations=6)
    res = d_res.download(np.int32, n)
    got = d_bits.download(np.uint8, n * mb)
    assert (res == 1).all(), res
    assert np.array_equal(got, np.concatenate(msgs))
    ctx.close()

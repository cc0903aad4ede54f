"""GPU TB reassembly + TB CRC (lh_tb_finalize) against the oracle's concatenation
(pusch_decoder_impl.cpp:612-781): directly from codeblock bits with chosen codeblock
verdicts, and end to end after lh_pusch_decode_batch."""
import numpy as np
import pytest

import ocudu_hip as H
import oracle as O
from gpu_util import ROW_MODES, make_ctx, require_gpu
from tb_util import cb_data_bits, tb_codeblocks, tb_crc_poly

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=ROW_MODES)
def ctx(request):
    c = make_ctx(4096, request.param)
    yield c
    c.close()


def _tb(bg, tbs, qm, nl, ns, seed):
    tb = np.random.default_rng(seed).integers(0, 256, tbs, dtype=np.uint8)
    _, seg = O.segment_rx(bg, tbs, qm, nl, ns)
    return tb, seg, tb_codeblocks(tb, seg)


def test_tb_finalize_direct(ctx):
    print("case tb_finalize direct", flush=True)
    # (bg, TB bytes, Qm, layers, symbols, verdict): "ok", "cbfail" (one CB CRC failed),
    # "tbfail" (a data bit flipped with every CB CRC passing), "ckfail" (checksum flipped).
    cases = [
        (2, 38, 4, 1, 600, "ok"), (2, 38, 4, 1, 600, "cbfail"), (1, 479, 2, 1, 4000, "ok"),
        (1, 3752, 2, 1, 40002, "ok"), (1, 3752, 2, 1, 40002, "tbfail"), (1, 3752, 2, 1, 40002, "cbfail"),
        (1, 20000, 8, 4, 96000, "ok"), (1, 20000, 8, 4, 96000, "ckfail"), (2, 2000, 2, 1, 17001, "ok"),
        (1, 8433, 6, 3, 36003, "tbfail"), (1, 159744, 8, 4, 170000, "ok"), (1, 6250, 8, 2, 26000, "ok"),
    ]
    rng = np.random.default_rng(3)
    descs, expect, parts, results = [], [], [], []
    cb_off = tb_off = first = 0
    for i, (bg, tbs, qm, nl, ns, verdict) in enumerate(cases):
        tb, seg, bits = _tb(bg, tbs, qm, nl, ns, seed=100 + i)
        C, K = bits.shape
        D = cb_data_bits(seg)
        res = np.full(C, 3, np.int32)
        if verdict == "cbfail":
            res[int(rng.integers(0, C))] = 0
        elif verdict == "tbfail":
            c = int(rng.integers(0, C))
            bits[c, int(rng.integers(0, min(D, 8 * tbs - c * D)))] ^= 1
        elif verdict == "ckfail":
            bits[C - 1, 8 * tbs - (C - 1) * D + int(rng.integers(0, 24))] ^= 1
        stride = (K // 8 + 15) // 16 * 16
        packed = np.zeros((C, stride), np.uint8)
        for c in range(C):
            packed[c, :K // 8] = np.packbits(bits[c])
        parts.append((cb_off, packed.reshape(-1)))
        results.append(res)
        descs.append(H.TbDesc(tb_off, cb_off, stride, 8 * tbs, first, C, D, 0))
        if C == 1:
            ok = res[0] > 0
            exp_status = 1 if ok else 0
            exp_tb = np.packbits(bits[0])[:tbs] if ok else None
        elif (res <= 0).any():
            exp_status, exp_tb = 0, None
        else:
            got, r = O.tb_concat_crc(packed.reshape(-1), stride, tbs, seg)
            exp_status, exp_tb = (1 if r == 0 else -1), got
        if verdict == "ok":
            assert exp_status == 1 and np.array_equal(exp_tb, tb)
        expect.append((tb_off, tbs, exp_status, exp_tb))
        cb_off += packed.size
        tb_off += (tbs + 15) // 16 * 16
        first += C
    cbb = np.zeros(cb_off, np.uint8)
    for o, a in parts:
        cbb[o:o + a.size] = a
    d_cb = H.DeviceBuffer.from_array(cbb)
    d_res = H.DeviceBuffer.from_array(np.concatenate(results))
    sentinel = np.full(tb_off, 0xA5, np.uint8)
    d_tb = H.DeviceBuffer.from_array(sentinel)
    d_st = H.DeviceBuffer(4 * len(descs))
    ctx.tb_finalize(descs, d_cb, d_res, d_tb, d_st)
    st = d_st.download(np.int32, len(descs))
    out = d_tb.download(np.uint8, tb_off)
    for i, (o, tbs, es, et) in enumerate(expect):
        assert st[i] == es, f"tb {i} ({cases[i]}): status {st[i]} vs {es}"
        if et is not None:
            assert np.array_equal(out[o:o + tbs], et), f"tb {i}: bytes differ"
        elif es == 0:
            assert (out[o:o + tbs] == 0xA5).all(), f"tb {i}: written although a codeblock failed"


def test_tb_finalize_rejects_bad_sizes(ctx):
    d = H.DeviceBuffer(64)
    bad = [H.TbDesc(0, 0, 16, 12, 0, 1, 0, 0),            # not a whole number of bytes
           H.TbDesc(0, 0, 1056, 8 * 3000, 0, 3, 8000, 0),  # data bits do not cover TB + CRC
           H.TbDesc(0, 0, 1056, 800, 0, 3, 8000, 0)]       # TB ends before the last codeblock
    for b in bad:
        with pytest.raises(H.LhError):
            ctx.tb_finalize([b], d, d, d, d)


def test_tb_decode_and_finalize(ctx):
    print("case tb decode + finalize", flush=True)
    # Two TBs in one decoder batch and one finalize call: 4 CBs of BG1 and 1 CB of BG2.
    tbspec = [(1, 3752, 4, 1, 12000, 11), (2, 300, 2, 1, 2000, 12)]
    cbs, cw_parts, tbdescs, tbs_all, expect = [], [], [], [], []
    cw_off = rm_off = bits_off = first = tb_off = 0
    for bg, tbs, qm, nl, ns, seed in tbspec:
        tb, seg, bits = _tb(bg, tbs, qm, nl, ns, seed)
        md, _ = O.segment_rx(bg, tbs, qm, nl, ns)
        C, K = bits.shape
        Z, F = seg.Z, seg.nof_filler_bits
        N = O.N_short(bg) * Z
        stride = H.msg_bytes(bg, Z)
        cb_poly = O.CRC24B if C > 1 else tb_crc_poly(8 * tbs)
        for c in range(C):
            E = md[c]["rm_length"]
            cw = O.encode(bits[c], bg, Z)
            e = O.rate_match(cw[2 * Z:], E, bg, Z, rv=0, Qm=qm, F=F)
            llr = O.awgn_llrs(e, qm, 12.0, seed * 1000 + c)
            rm = O.rate_dematch(llr, bg, Z, rv=0, Qm=qm, F=F)
            expect.append(O.decode(rm, bg, Z, F=F, crc_poly=cb_poly, max_iter=6, crc_bits=md[c]["nof_crc_bits"])[0])
            cbs.append(H.PuschCbDesc(H.RmDesc(cw_off, rm_off, E, bg, Z, 0, qm, 0, F), bits_off + c * stride,
                                     md[c]["nof_crc_bits"], cb_poly))
            cw_parts.append((cw_off, llr))
            cw_off += E
            rm_off += N
        tbdescs.append(H.TbDesc(tb_off, bits_off, stride, 8 * tbs, first, C, cb_data_bits(seg), 0))
        tbs_all.append((tb_off, tb))
        bits_off += C * stride
        first += C
        tb_off += tbs
    cw = np.zeros(cw_off, np.int8)
    for o, a in cw_parts:
        cw[o:o + a.size] = a
    d_cw = H.DeviceBuffer.from_array(cw)
    d_rm = H.DeviceBuffer(rm_off)
    d_bits = H.DeviceBuffer(bits_off)
    d_bits.zero()
    d_res = H.DeviceBuffer(4 * len(cbs))
    ctx.pusch_decode_batch(cbs, d_cw, d_rm, d_bits, d_res, new_data=True, use_early_stop=True,
                           nof_ldpc_iterations=6)
    d_tb = H.DeviceBuffer(tb_off)
    d_st = H.DeviceBuffer(4 * len(tbdescs))
    ctx.tb_finalize(tbdescs, d_bits, d_res, d_tb, d_st)
    res = d_res.download(np.int32, len(cbs))
    assert list(res) == expect and (res > 0).all(), (res, expect)
    st = d_st.download(np.int32, len(tbdescs))
    out = d_tb.download(np.uint8, tb_off)
    for i, (o, tb) in enumerate(tbs_all):
        assert st[i] == 1
        assert np.array_equal(out[o:o + tb.size], tb), f"tb {i}: bytes differ"

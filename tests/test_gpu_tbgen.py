"""Device TB generator (lh_gen_transport_blocks): TB bytes follow the splitmix64 stream,
codeblock messages equal the TX segmentation of those bytes (tb_util.tb_codeblocks,
pinned against the reference's TX segmenter), and the whole transmit chain (TB generator
-> codeblock generator with message input -> PUSCH decode -> TB finalize) returns the TB."""
import numpy as np
import pytest

import ocudu_hip as H
import oracle as O
from gpu_util import require_gpu
from tb_util import tb_codeblocks

pytestmark = pytest.mark.gpu

M64 = (1 << 64) - 1
GAMMA = 0x9e3779b97f4a7c15


def _mix(z):
    z = ((z ^ (z >> 30)) * 0xbf58476d1ce4e5b9) & M64
    z = ((z ^ (z >> 27)) * 0x94d049bb133111eb) & M64
    return z ^ (z >> 31)


def _tb_bytes(seed, n):
    out = np.zeros(n, np.uint8)
    for w in range((n + 7) // 8):
        x = _mix((seed + (w + 1) * GAMMA) & M64)
        for k in range(min(8, n - 8 * w)):
            out[8 * w + k] = (x >> (56 - 8 * k)) & 0xff
    return out


CASES = [(2, 38, 4, 1, 600), (1, 479, 2, 1, 4000), (1, 3752, 4, 1, 12000), (1, 20000, 8, 4, 96000),
         (2, 2000, 2, 1, 17001), (1, 8433, 6, 3, 36003)]


def test_tb_generator_and_chain():
    print("case TB generator chain", flush=True)
    require_gpu()
    ctx = H.Context(0, max_codeblocks=2048)
    descs, info, gens, pusch, tbds = [], [], [], [], []
    tb_off = msg_off = cw_off = rm_off = first = 0
    for i, (bg, tbs, qm, nl, ns) in enumerate(CASES):
        seg, rm, off = H.segment_rx(bg, 8 * tbs, qm, nl, ns)
        Z, C = seg.lifting_size, seg.nof_cbs
        stride = H.msg_bytes(bg, Z)
        seed = 1000 + 17 * i
        descs.append(H.GenTbDesc(tb_off, msg_off, seed, stride, 8 * tbs, C, seg.segment_length, seg.nof_filler_bits))
        poly = O.CRC24B if C > 1 else (O.CRC16 if 8 * tbs <= 3824 else O.CRC24A)
        snr = {2: 12.0, 4: 16.0, 6: 20.0, 8: 26.0}[qm]
        for c in range(C):
            gens.append(H.GenDesc(cw_off + off[c], msg_off + c * stride, 0, 77 + c, 10 ** (-snr / 10), rm[c], bg, Z, 0,
                                  qm, 0, seg.nof_filler_bits, poly, H.LH_GEN_MSG_INPUT, 0))
            pusch.append(H.PuschCbDesc(H.RmDesc(cw_off + off[c], rm_off, rm[c], bg, Z, 0, qm, 0, seg.nof_filler_bits),
                                       msg_off + c * stride, seg.nof_cb_crc_bits, poly))
            rm_off += O.N_short(bg) * Z
        tbds.append(H.TbDesc(tb_off, msg_off, stride, 8 * tbs, first, C, seg.cb_data_bits, 0))
        info.append((tb_off, msg_off, tbs, bg, qm, nl, ns, seed, stride, C))
        tb_off += tbs
        msg_off += C * stride
        cw_off += off[-1] + rm[-1]
        first += C
    d_tb = H.DeviceBuffer(tb_off)
    d_msg = H.DeviceBuffer(msg_off)
    ctx.gen_transport_blocks(descs, d_tb, d_msg)
    tb_all = d_tb.download(np.uint8, tb_off)
    msg_all = d_msg.download(np.uint8, msg_off)
    for to, mo, tbs, bg, qm, nl, ns, seed, stride, C in info:
        tb = tb_all[to:to + tbs]
        assert np.array_equal(tb, _tb_bytes(seed, tbs)), f"TB {tbs}: bytes"
        _, oseg = O.segment_rx(bg, tbs, qm, nl, ns)
        exp = tb_codeblocks(tb, oseg)
        for c in range(C):
            got = np.unpackbits(msg_all[mo + c * stride:mo + (c + 1) * stride])[:exp.shape[1]]
            assert np.array_equal(got, exp[c]), f"TB {tbs}: codeblock {c} message"
    # Chain: encode the generated messages, decode, reassemble.
    d_cw = H.DeviceBuffer(cw_off)
    ctx.gen_codeblocks(gens, d_cw, d_msg)
    d_rm = H.DeviceBuffer(rm_off)
    d_bits = H.DeviceBuffer(msg_off)
    d_bits.zero()
    d_res = H.DeviceBuffer(4 * len(pusch))
    ctx.pusch_decode_batch(pusch, d_cw, d_rm, d_bits, d_res, nof_ldpc_iterations=8)
    d_tb2 = H.DeviceBuffer(tb_off)
    d_st = H.DeviceBuffer(4 * len(tbds))
    ctx.tb_finalize(tbds, d_bits, d_res, d_tb2, d_st)
    assert (d_res.download(np.int32, len(pusch)) > 0).all()
    assert (d_st.download(np.int32, len(tbds)) == 1).all()
    assert np.array_equal(d_tb2.download(np.uint8, tb_off), tb_all)
    ctx.close()

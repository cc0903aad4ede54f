"""Transport-block test data: TX segmentation of a TB into codeblock messages (TS38.212
5.1-5.2 as done by lib/phy/upper/channel_coding/ldpc/ldpc_segmenter_tx_impl.cpp:39-198),
built on the oracle's CRC and segmentation.  Checked against the reference TX segmenter
in test_oracle_segmentation.py."""
import numpy as np

import oracle as O

MAX_BITS_CRC16 = 3824


def tb_crc_poly(tbs_bits):
    return O.CRC16 if tbs_bits <= MAX_BITS_CRC16 else O.CRC24A


def _with_crc(bits, poly, order):
    r = O.crc(poly, np.packbits(bits), bits.size)
    tail = ((r >> np.arange(order - 1, -1, -1)) & 1).astype(np.uint8)
    return np.concatenate([bits, tail])


def tb_codeblocks(tb_bytes, seg):
    """Codeblock messages (C, K_b*Z) uint8 of a TB: data, TB CRC, zero padding, CB CRC24B
    when C > 1, zero fillers.  seg: the oracle segmentation (oracle.segment_rx)."""
    tb_bytes = np.asarray(tb_bytes, np.uint8)
    tbs = 8 * tb_bytes.size
    C, K, F = seg.nof_segments, seg.segment_length, seg.nof_filler_bits
    tbcrc = 16 if tbs <= MAX_BITS_CRC16 else 24
    b = _with_crc(np.unpackbits(tb_bytes), tb_crc_poly(tbs), tbcrc)
    out = np.zeros((C, K), np.uint8)
    if C == 1:
        out[0, :b.size] = b
        assert b.size == K - F
        return out
    D = K - F - 24
    padded = np.zeros(C * D, np.uint8)
    padded[:b.size] = b
    for c in range(C):
        out[c, :K - F] = _with_crc(padded[c * D:(c + 1) * D], O.CRC24B, 24)
    return out


def cb_data_bits(seg):
    """Data bits per codeblock as pusch_decoder_impl::concatenate_codeblocks counts them."""
    return seg.segment_length - seg.nof_filler_bits - (24 if seg.nof_segments > 1 else 0)

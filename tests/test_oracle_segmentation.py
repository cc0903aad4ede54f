"""Oracle segmentation and TB reassembly pinned against the reference's own segmenters
(ldpc_segmenter_rx_impl.cpp / ldpc_segmenter_tx_impl.cpp, built from source into
oracle/_ref).  TB reassembly is the inverse of TX segmentation, so the reference TX
segmenter's codeblocks must reassemble into the original TB with a matching TB CRC,
and a flipped bit anywhere in the data or the checksum must break it."""
import numpy as np
import pytest

import oracle as O
from tb_util import cb_data_bits, tb_codeblocks

# (bg, TB bytes, Qm, layers, channel symbols): C = 1 with CRC16 and CRC24A, several
# codeblocks with and without zero padding, short/long rate-matched lengths, BG2.
CASES = [
    (2, 4, 2, 1, 120), (2, 38, 4, 1, 600), (2, 478, 2, 2, 4000), (1, 479, 2, 1, 4000),
    (1, 1000, 6, 1, 3000), (1, 3752, 2, 1, 40002), (1, 6250, 8, 2, 26000), (2, 600, 4, 1, 3333),
    (1, 20000, 8, 4, 96000), (1, 159744, 8, 4, 170000), (2, 2000, 2, 1, 17001), (1, 8433, 6, 3, 36003),
]


@pytest.fixture(scope="module")
def ref():
    R = O.ref()
    if R is None:
        pytest.skip("reference build oracle/_ref not present")
    return R


@pytest.mark.parametrize("case", CASES)
def test_segment_rx_matches_reference(ref, case):
    bg, tbs, qm, nl, ns = case
    mo, _ = O.segment_rx(bg, tbs, qm, nl, ns)
    mr = O.ref_segment_rx(bg, tbs, qm, nl, ns)
    mt, _ = O.ref_segment_tx(bg, np.zeros(tbs, np.uint8), qm, nl, ns, with_bits=False)
    assert mo == mr
    assert [dict(m, info_bits=0) for m in mt] == mr


@pytest.mark.parametrize("case", CASES)
def test_tb_builder_matches_reference_tx(ref, case):
    bg, tbs, qm, nl, ns = case
    tb = np.random.default_rng(tbs).integers(0, 256, tbs, dtype=np.uint8)
    md, bits = O.ref_segment_tx(bg, tb, qm, nl, ns)
    _, seg = O.segment_rx(bg, tbs, qm, nl, ns)
    assert np.array_equal(tb_codeblocks(tb, seg), bits)


@pytest.mark.parametrize("case", CASES)
def test_tb_concat_inverts_reference_tx(ref, case):
    bg, tbs, qm, nl, ns = case
    rng = np.random.default_rng(7 + tbs)
    tb = rng.integers(0, 256, tbs, dtype=np.uint8)
    md, bits = O.ref_segment_tx(bg, tb, qm, nl, ns)
    _, seg = O.segment_rx(bg, tbs, qm, nl, ns)
    C, K = len(md), md[0]["segment_length"]
    stride = K // 8
    packed = np.stack([np.packbits(b) for b in bits])
    if C == 1:
        # One codeblock: the TB CRC is the codeblock CRC over data + checksum.
        poly = O.CRC16 if 8 * tbs <= 3824 else O.CRC24A
        assert O.crc(poly, packed[0], K - seg.nof_filler_bits) == 0
        return
    for c in range(C):
        assert O.crc(O.CRC24B, packed[c], K - seg.nof_filler_bits) == 0
    got, r = O.tb_concat_crc(packed.reshape(-1), stride, tbs, seg)
    assert np.array_equal(got, tb) and r == 0
    D = cb_data_bits(seg)
    # A flipped data bit in a random codeblock, then a flipped checksum bit.
    for c, j in [(int(rng.integers(0, C)), None), (C - 1, 8 * tbs - (C - 1) * D + int(rng.integers(0, 24)))]:
        if j is None:
            j = int(rng.integers(0, min(D, 8 * tbs - c * D)))
        bad = bits.copy()
        bad[c, j] ^= 1
        pb = np.stack([np.packbits(b) for b in bad])
        _, r = O.tb_concat_crc(pb.reshape(-1), stride, tbs, seg)
        assert r != 0, (c, j)


def test_product_segmentation_matches_oracle():
    """lh_segment_rx (the C-ABI's host segmentation) against the oracle's restatement over
    a sweep of sizes, layers and modulations (the oracle is pinned to the reference above)."""
    import ocudu_hip as H
    rng = np.random.default_rng(17)
    cases = list(CASES)
    for _ in range(300):
        bg = int(rng.integers(1, 3))
        tbs = int(rng.integers(1, 159000))
        qm = int(rng.choice([2, 4, 6, 8]))
        nl = int(rng.integers(1, 5))
        ns = nl * int(rng.integers(1, 60000))
        cases.append((bg, tbs, qm, nl, ns))
    for bg, tbs, qm, nl, ns in cases:
        seg, rm, off = H.segment_rx(bg, 8 * tbs, qm, nl, ns)
        if seg.nof_cbs > 256:  # beyond the oracle's arrays (BG2 at the largest sizes)
            continue
        mo, oseg = O.segment_rx(bg, tbs, qm, nl, ns)
        assert seg.nof_cbs == len(mo) and seg.lifting_size == oseg.Z, (bg, tbs)
        assert seg.segment_length == oseg.segment_length and seg.nof_filler_bits == oseg.nof_filler_bits
        assert seg.nof_cb_crc_bits == oseg.nof_crc_bits and seg.nof_tb_crc_bits == oseg.nof_tb_crc_bits
        assert rm == [m["rm_length"] for m in mo] and off == [m["cw_offset"] for m in mo]
        if seg.nof_cbs > 1:
            assert seg.cb_data_bits == cb_data_bits(oseg)

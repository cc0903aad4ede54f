"""Slot graphs (lh_slot_graph_*): a captured slot replayed several times, with eager work
and synchronisations between the replays, must reproduce the eager calls exactly on
every replay -- decoder results, message bits, HARQ buffers, TB bytes and TB verdicts --
also after the codeword LLRs change between replays, and for the learned decoder with
its plain fallback."""
from pathlib import Path

import numpy as np
import pytest

import ocudu_hip as H
import oracle as O
from gpu_util import ROW_MODES, make_ctx, require_gpu
from tb_util import cb_data_bits, tb_codeblocks, tb_crc_poly

pytestmark = pytest.mark.gpu
TABLES = Path(__file__).resolve().parent.parent / "ocudu-dapp-platform_amd" / "tables"


@pytest.fixture(scope="module", params=ROW_MODES)
def ctx(request):
    c = make_ctx(1024, request.param)
    yield c
    c.close()


def _slot(snr_shift=0.0, noise_seed=0):
    """Three TBs (4 + 1 + 6 codeblocks, mixed BG/Z/Qm) as one slot."""
    spec = [(1, 3752, 4, 1, 12000, 11, 12.0), (2, 300, 2, 1, 2000, 12, 6.0), (1, 6250, 6, 2, 26000, 13, 17.0)]
    cbs, parts, tbd = [], [], []
    cw = rm = bo = first = to = 0
    for bg, tbs, qm, nl, ns, seed, snr in spec:
        tb = np.random.default_rng(seed).integers(0, 256, tbs, dtype=np.uint8)
        md, seg = O.segment_rx(bg, tbs, qm, nl, ns)
        bits = tb_codeblocks(tb, seg)
        C = bits.shape[0]
        Z, F = seg.Z, seg.nof_filler_bits
        stride = H.msg_bytes(bg, Z)
        poly = O.CRC24B if C > 1 else tb_crc_poly(8 * tbs)
        for c in range(C):
            E = md[c]["rm_length"]
            e = O.rate_match(O.encode(bits[c], bg, Z)[2 * Z:], E, bg, Z, rv=0, Qm=qm, F=F)
            parts.append((cw, O.awgn_llrs(e, qm, snr + snr_shift, seed * 100 + c + noise_seed)))
            cbs.append(H.PuschCbDesc(H.RmDesc(cw, rm, E, bg, Z, 0, qm, 0, F), bo + c * stride, md[c]["nof_crc_bits"],
                                     poly))
            cw += E
            rm += O.N_short(bg) * Z
        tbd.append(H.TbDesc(to, bo, stride, 8 * tbs, first, C, cb_data_bits(seg), 0))
        bo += C * stride
        first += C
        to += tbs
    llr = np.zeros(cw, np.int8)
    for o, a in parts:
        llr[o:o + a.size] = a
    return cbs, tbd, llr, rm, bo, to


class _Bufs:
    def __init__(self, cw, rm, bits, n, tb, ntb):
        self.d_cw, self.d_rm, self.d_bits = H.DeviceBuffer(cw), H.DeviceBuffer(rm), H.DeviceBuffer(bits)
        self.d_res, self.d_tb, self.d_st = H.DeviceBuffer(4 * n), H.DeviceBuffer(tb), H.DeviceBuffer(4 * ntb)
        self.sizes = (rm, bits, n, tb, ntb)

    def zero(self):
        for b in (self.d_rm, self.d_bits, self.d_res, self.d_tb, self.d_st):
            b.zero()

    def snapshot(self):
        rm, bits, n, tb, ntb = self.sizes
        return (self.d_rm.download(np.int8, rm), self.d_bits.download(np.uint8, bits),
                self.d_res.download(np.int32, n), self.d_tb.download(np.uint8, tb), self.d_st.download(np.int32, ntb))


def _eager(ctx, B, cbs, tbd, learned):
    B.zero()
    if learned:
        ctx.pusch_decode_batch_learned(cbs, B.d_cw, B.d_rm, B.d_bits, B.d_res, nof_ldpc_iterations=8,
                                       nof_fallback_iterations=8)
    else:
        ctx.pusch_decode_batch(cbs, B.d_cw, B.d_rm, B.d_bits, B.d_res, nof_ldpc_iterations=8)
    ctx.tb_finalize(tbd, B.d_bits, B.d_res, B.d_tb, B.d_st)
    H.hip().hipDeviceSynchronize()
    return B.snapshot()


@pytest.mark.parametrize("learned", [False, True])
def test_slot_graph_replays_match_eager(ctx, learned):
    print(f"case slot graph learned={learned}", flush=True)
    if learned:
        for bg in (1, 2):
            from ocudu_hip import nms
            _, aq, bq, _ = nms.load_table(TABLES / f"nms_bg{bg}.json")
            ctx.set_scale_table(bg, aq, bq)
    cbs, tbd, llr, rmb, bb, tbb = _slot()
    B = _Bufs(llr.size, rmb, bb, len(cbs), tbb, len(tbd))
    other = _Bufs(llr.size, rmb, bb, len(cbs), tbb, len(tbd))
    B.d_cw.upload(llr)
    other.d_cw.upload(llr[::-1].copy())
    ref = _eager(ctx, B, cbs, tbd, learned)
    assert (ref[4] == 1).all(), ref[4]
    g = ctx.slot_graph(cbs, B.d_cw, B.d_rm, B.d_bits, B.d_res, nof_ldpc_iterations=8, learned=learned,
                       nof_fallback_iterations=8 if learned else 0, tbs=tbd, d_tb=B.d_tb, d_tb_status=B.d_st)
    for rep in range(4):
        B.zero()
        g.launch()
        H.hip().hipDeviceSynchronize()
        got = B.snapshot()
        for k, (x, y) in enumerate(zip(ref, got)):
            assert np.array_equal(x, y), f"replay {rep}: output {k} differs"
        _eager(ctx, other, cbs, tbd, learned)  # unrelated eager work between replays
    # New LLRs in place (lower SNR, other noise): the replay decodes them like an eager call.
    _, _, llr2, *_ = _slot(snr_shift=-3.0, noise_seed=7)
    B.d_cw.upload(llr2)
    ref2 = _eager(ctx, B, cbs, tbd, learned)
    for rep in range(2):
        B.zero()
        g.launch()
        H.hip().hipDeviceSynchronize()
        got = B.snapshot()
        for k, (x, y) in enumerate(zip(ref2, got)):
            assert np.array_equal(x, y), f"new LLRs, replay {rep}: output {k} differs"
    g.close()

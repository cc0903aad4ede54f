"""Serving runtime (lh_server_*): every transport block's result -- TB CRC verdict,
codeblocks passed, iterations, TB bytes -- is the same when it is served alone as when it
is served in the mix of a continuous-batching run with overlapping slots, HARQ
retransmissions (rv2 soft-combined into device buffers) included, for the plain and the
learned decoder.  Bit-exact (tolerance zero)."""
from pathlib import Path

import numpy as np
import pytest

import ocudu_hip as H
from gpu_util import require_gpu

pytestmark = pytest.mark.gpu
TABLES = Path(__file__).resolve().parent.parent / "ocudu-dapp-platform_amd" / "tables"


@pytest.fixture(scope="module")
def setup():
    require_gpu()
    from ocudu_hip import serve
    ctx = H.Context(0, max_codeblocks=4096)
    thr = serve.mcs_thresholds(ctx, n=64)
    trace = serve.Trace(ctx, slots=12, mean_ues=8, thresholds=thr, seed=3, snr_lo=-0.6, snr_hi=0.4, prb_max=48)
    yield ctx, trace
    ctx.close()


def _alone(ctx, trace, i, learned, mix):
    from ocudu_hip import serve
    it = serve.PUSCH_LDPC_ITERATIONS  # as in the mix
    srv = H.Server(ctx, 1 << 28, 1024, 1, it, learned=learned, nof_fallback_iterations=it)
    out = {}
    for tx in (0, 1):
        if 2 * i + tx not in mix:
            break
        buf = np.zeros(trace.tbs[i]["tbs"] // 8, np.uint8)
        srv.submit(trace.request(i, tx), buf.ctypes.data)
        srv.step()
        got = []
        while not got:
            got = srv.poll(wait=True)
        out[2 * i + tx] = (got[0], buf)
    srv.close()
    return out


@pytest.mark.parametrize("learned", [False, True])
def test_serving_alone_equals_mix(setup, learned):
    print(f"case serving alone vs mix learned={learned}", flush=True)
    from ocudu_hip import nms, serve
    ctx, trace = setup
    if learned:
        for bg in (1, 2):
            _, aq, bq, _ = nms.load_table(TABLES / f"nms_bg{bg}.json")
            ctx.set_scale_table(bg, aq, bq)
    mix, slot_lat, wall, bufs = serve.serve(ctx, trace, learned=learned, realtime=False, slots_in_flight=3,
                                           timeout_s=120)
    assert len(mix) >= len(trace.tbs)
    n_retx = sum(1 for k in mix if k % 2 == 1)
    assert n_retx > 0, "the trace should exercise HARQ retransmissions"
    for i in range(len(trace.tbs)):
        alone = _alone(ctx, trace, i, learned, mix)
        for key, (r, buf) in alone.items():
            m = mix[key]
            assert (r.tb_crc_ok, r.nof_cbs_ok, r.max_iterations) == (m.tb_crc_ok, m.nof_cbs_ok, m.max_iterations), \
                f"TB {i} transmission {key % 2}: alone {r.tb_crc_ok, r.nof_cbs_ok, r.max_iterations} vs mix " \
                f"{m.tb_crc_ok, m.nof_cbs_ok, m.max_iterations}"
            if r.tb_crc_ok == H.LH_TB_CRC_OK:
                assert np.array_equal(buf, bufs[key])
                t = trace.tbs[i]
                assert np.array_equal(buf, trace.tb_host[t["tb_off"]:t["tb_off"] + t["tbs"] // 8])
    s = serve.summarize(trace, mix, slot_lat, wall, bufs)
    print(s, flush=True)
    assert s["wrong_tbs"] == 0


def test_admission_waits_for_memory(setup):
    """A HARQ budget that holds one TB at a time: every request is still served, in order,
    as completions release memory."""
    print("case admission by budget", flush=True)
    ctx, trace = setup
    from ocudu_hip import serve
    # A HARQ process holds per codeblock N_short*Z soft bits and its message bits.
    need = max(t["seg"].nof_cbs * ((66 if t["bg"] == 1 else 50) * t["seg"].lifting_size
                                   + ((22 if t["bg"] == 1 else 10) * t["seg"].lifting_size + 7) // 8)
               for t in trace.tbs)
    mix, slot_lat, wall, bufs = serve.serve(ctx, trace, realtime=False, harq_budget=need + 4096, timeout_s=120)
    assert all(2 * i in mix for i in range(len(trace.tbs)))
    assert all(r.tb_crc_ok != H.LH_TB_REJECTED for r in mix.values())


def test_servers_recreated(setup):
    """Servers created and destroyed one after another on one context (plain, learned,
    plain, the bench's order): every run completes without a runtime error, and the two
    plain runs give the same result for every transmission."""
    print("case servers recreated", flush=True)
    from ocudu_hip import nms, serve
    ctx, trace = setup
    for bg in (1, 2):
        _, aq, bq, _ = nms.load_table(TABLES / f"nms_bg{bg}.json")
        ctx.set_scale_table(bg, aq, bq)
    runs = [serve.serve(ctx, trace, learned=learned, realtime=False, timeout_s=120)[0]
            for learned in (False, True, False)]
    a, b = runs[0], runs[2]
    assert set(a) == set(b)
    for k in a:
        assert (a[k].tb_crc_ok, a[k].nof_cbs_ok, a[k].max_iterations) == \
            (b[k].tb_crc_ok, b[k].nof_cbs_ok, b[k].max_iterations), k


class _At:
    """A device address (the C ABI only reads .ptr)."""

    def __init__(self, ptr):
        self.ptr = ptr


def test_passed_codeblocks_are_not_decoded_again(setup):
    """HARQ codeblock CRC state (pusch_decoder_impl.cpp:412-431): a codeblock whose CRC
    passed in the first transmission keeps its bits and is not decoded again.  The second
    transmission's LLRs of those codeblocks are replaced by noise: the server's TB results
    must equal a model that decodes only the failed codeblocks (with HARQ combining) and
    keeps the others, and the TBs that this completes must reassemble exactly."""
    print("case HARQ codeblock CRC state", flush=True)
    from ocudu_hip import serve
    ctx, _ = setup
    it = serve.PUSCH_LDPC_ITERATIONS
    thr = serve.mcs_thresholds(ctx, n=64)
    trace = serve.Trace(ctx, slots=8, mean_ues=8, thresholds=thr, seed=17, snr_lo=-0.5, snr_hi=-0.2, prb_max=120)
    multi = [i for i, t in enumerate(trace.tbs) if t["seg"].nof_cbs >= 2]
    assert multi, "the trace should hold multi-codeblock TBs"

    def cb_descs(i, tx, rm_base):
        t = trace.tbs[i]
        seg = t["seg"]
        n = (66 if t["bg"] == 1 else 50) * seg.lifting_size
        mb = t["stride"]
        return [H.PuschCbDesc(H.RmDesc(t["off"][c], rm_base + c * n, t["rm"][c], t["bg"], seg.lifting_size, 2 * tx,
                                       t["qm"], 0, seg.nof_filler_bits), c * mb, 24, H.LH_CRC24B)
                for c in range(seg.nof_cbs)]

    # Model, first transmission: which codeblocks pass.
    model = {}
    for i in multi:
        t = trace.tbs[i]
        C, mb = t["seg"].nof_cbs, t["stride"]
        d_rm = H.DeviceBuffer(C * (66 if t["bg"] == 1 else 50) * t["seg"].lifting_size)
        d_bits = H.DeviceBuffer(C * mb)
        d_res = H.DeviceBuffer(4 * C)
        ctx.pusch_decode_batch(cb_descs(i, 0, 0), _At(trace.d_llr.ptr + t["llr_off"][0]), d_rm, d_bits, d_res,
                               new_data=True, nof_ldpc_iterations=it)
        model[i] = (d_rm, d_bits, d_res.download(np.int32, C))
    cand = [i for i in multi if 0 < (model[i][2] > 0).sum() < trace.tbs[i]["seg"].nof_cbs]
    assert cand, "no TB with some codeblocks passing in its first transmission"
    # Noise in place of the second transmission's LLRs of the codeblocks that passed.
    rng = np.random.default_rng(3)
    for i in cand:
        t = trace.tbs[i]
        for c in np.flatnonzero(model[i][2] > 0):
            noise = rng.choice(np.array([-60, 60], np.int8), t["rm"][c])
            trace.d_llr.upload(noise, offset=t["llr_off"][1] + t["off"][c])
    # Model, second transmission: the failed codeblocks decode with combining.
    expect = {}
    for i in cand:
        t = trace.tbs[i]
        C, mb = t["seg"].nof_cbs, t["stride"]
        d_rm, d_bits0, r0 = model[i]
        d_bits = H.DeviceBuffer(C * mb)
        d_res = H.DeviceBuffer(4 * C)
        ctx.pusch_decode_batch(cb_descs(i, 1, 0), _At(trace.d_llr.ptr + t["llr_off"][1]), d_rm, d_bits, d_res,
                               new_data=False, nof_ldpc_iterations=it)
        r2 = d_res.download(np.int32, C)
        expect[i] = bool(((r0 > 0) | (r2 > 0)).all())
    assert any(expect.values()), "no TB completed by its retransmission"
    # The server: first transmissions, then the second ones.
    srv = H.Server(ctx, 1 << 30, 4096, 1, it, learned=False, nof_fallback_iterations=it)
    bufs = {}
    for tx in (0, 1):
        for i in cand:
            bufs[(i, tx)] = np.zeros(trace.tbs[i]["tbs"] // 8, np.uint8)
            srv.submit(trace.request(i, tx), bufs[(i, tx)].ctypes.data)
        srv.step()
        got = {}
        while len(got) < len(cand):
            for r in srv.poll(wait=True):
                got[r.id // 2] = r
        if tx == 1:
            for i in cand:
                ok = got[i].tb_crc_ok == H.LH_TB_CRC_OK
                assert ok == expect[i], (i, got[i].tb_crc_ok, expect[i])
                if ok:
                    t = trace.tbs[i]
                    assert np.array_equal(bufs[(i, 1)], trace.tb_host[t["tb_off"]:t["tb_off"] + t["tbs"] // 8])
    srv.close()

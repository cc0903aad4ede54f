"""Serving runtime (lh_server_*).  Every transmission of a continuous-batching run (HARQ
rv0 + rv2 soft-combined on the device, codeblock CRC state, overlapping slots) equals the
oracle chain of its HARQ process (tests/serve_oracle.py): TB CRC verdict, codeblocks passed,
iterations and TB bytes; every transport block's result is the same served alone as in the
mix, for the plain and the learned decoder; the codeblock CRC state (passed codeblocks not
decoded again, reset after a TB CRC failure), admission by HBM budget, request validation
and HARQ expiry.  Bit-exact (tolerance zero)."""
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np
import pytest

import ocudu_hip as H
import oracle as O
from gpu_util import require_gpu
from serve_oracle import HarqModel
from tb_util import tb_codeblocks

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


def _model(t):
    return HarqModel(t["bg"], t["tbs"], t["qm"], t["layers"], t["nsym"])


def _same(r, want, what):
    got = (r.tb_crc_ok, r.nof_cbs_ok, r.max_iterations)
    assert got == tuple(want[:3]), f"{what}: server {got} vs oracle {tuple(want[:3])}"


def test_serving_matches_oracle(setup):
    """Every transmission of a deterministic continuous-batching run of the trace (rv0, and
    rv2 combined into the device HARQ buffers for the failed ones) against the oracle chain
    of its HARQ process."""
    print("case serving vs oracle", flush=True)
    from ocudu_hip import serve
    ctx, trace = setup
    it = serve.PUSCH_LDPC_ITERATIONS
    mix, _, _, bufs = serve.serve(ctx, trace, realtime=False, slots_in_flight=3, timeout_s=120)
    llr = trace.d_llr.download(np.int8, trace.d_llr.nbytes)

    def model(i):
        t = trace.tbs[i]
        m = _model(t)
        n = t["nsym"] * t["qm"]
        out = [m.transmit(llr[t["llr_off"][0]:t["llr_off"][0] + n], 0, True, it)]
        if 2 * i + 1 in mix:
            out.append(m.transmit(llr[t["llr_off"][1]:t["llr_off"][1] + n], 2, False, it))
        return out

    with ThreadPoolExecutor(8) as ex:
        want = list(ex.map(model, range(len(trace.tbs))))
    kinds = set()
    for i, w in enumerate(want):
        for tx, exp in enumerate(w):
            r = mix[2 * i + tx]
            _same(r, exp, f"TB {i} transmission {tx}")
            kinds.add((tx, exp[0]))
            if exp[0] == H.LH_TB_CRC_OK:
                assert np.array_equal(bufs[2 * i + tx], exp[3]), f"TB {i} transmission {tx}: TB bytes differ"
    assert (1, H.LH_TB_CRC_OK) in kinds and (0, H.LH_TB_CB_FAILED) in kinds, kinds


def _tb_llrs(ctx, model, cb_bits, rv, n0, seed):
    """Device LLRs of one transmission of a TB whose codeblock messages (C, K) are given."""
    C, stride = model.C, model.stride
    d_msg = H.DeviceBuffer.from_array(np.packbits(cb_bits, axis=1).reshape(-1))
    n = sum(c["rm_length"] for c in model.cbs)
    d_llr = H.DeviceBuffer(n + 256)
    gens = [H.GenDesc(c["cw_offset"], k * stride, 0, seed + k, n0, c["rm_length"], model.bg, model.Z, rv, model.qm,
                      0, model.F, model.poly, H.LH_GEN_MSG_INPUT, 0) for k, c in enumerate(model.cbs)]
    ctx.gen_codeblocks(gens, d_llr, d_msg)
    H.hip().hipDeviceSynchronize()
    return d_llr, d_llr.download(np.int8, n)


def _request(d_llr, rid, proc, model, layers, nsym, tx, slot=0):
    return H.TbRequest(id=rid, d_llrs=d_llr.ptr, harq_process=proc, new_data=int(tx == 0), base_graph=model.bg,
                       tbs_bits=model.tbs, modulation_order=model.qm, nof_layers=layers, nof_ch_symbols=nsym,
                       rv=2 * tx, Nref=0, slot=slot)


def _one(srv, req, tb_out=None):
    srv.submit(req, tb_out.ctypes.data if tb_out is not None else None)
    srv.step()
    got = []
    while not got:
        got = srv.poll(wait=True)
    return got[0]


def test_tb_crc_failure_resets_codeblock_state(setup):
    """A TB whose codeblock CRCs all pass but whose TB CRC fails (one data bit flipped and
    that codeblock's CRC24B recomputed): LH_TB_CRC_FAILED, and the retransmission decodes
    every codeblock again (pusch_decoder_impl.cpp:657-669) -- here from soft bits combined
    with a retransmission of other data, so that keeping the passed codeblocks would give
    another verdict than decoding them again."""
    print("case TB CRC failure", flush=True)
    from ocudu_hip import serve
    ctx, _ = setup
    it = serve.PUSCH_LDPC_ITERATIONS
    prb, qm, r1024, layers = 60, 4, 616, 2
    nsym = serve.RE_PER_PRB * prb * layers
    tbs = serve.tbs_of(prb, qm, r1024, layers)
    m = HarqModel(1, tbs, qm, layers, nsym)
    assert m.C > 1
    rng = np.random.default_rng(9)
    bits = tb_codeblocks(rng.integers(0, 256, tbs // 8, dtype=np.uint8), m.seg)
    k_data = m.K - m.F - 24
    bits[0, 5] ^= 1
    crc = O.crc(O.CRC24B, np.packbits(bits[0, :k_data]), k_data)
    bits[0, k_data:k_data + 24] = (crc >> np.arange(23, -1, -1)) & 1
    other = tb_codeblocks(rng.integers(0, 256, tbs // 8, dtype=np.uint8), m.seg)
    d0, l0 = _tb_llrs(ctx, m, bits, 0, 0.0, 100)
    d1, l1 = _tb_llrs(ctx, m, other, 2, 0.0, 200)
    srv = H.Server(ctx, 1 << 26, 1024, 1, it)
    w0 = m.transmit(l0, 0, True, it)
    assert w0[:2] == (H.LH_TB_CRC_FAILED, m.C)
    _same(_one(srv, _request(d0, 0, 7, m, layers, nsym, 0)), w0, "first transmission")
    keep = (H.LH_TB_CRC_FAILED, m.C, w0[2])  # what keeping the passed codeblocks would give
    w1 = m.transmit(l1, 2, False, it)
    assert tuple(w1[:3]) != keep, "the retransmission does not tell the reset apart"
    _same(_one(srv, _request(d1, 1, 7, m, layers, nsym, 1)), w1, "retransmission")
    srv.close()


def test_passed_codeblocks_are_not_decoded_again(setup):
    """HARQ codeblock CRC state (pusch_decoder_impl.cpp:412-431): a codeblock whose CRC
    passed in the first transmission keeps its bits and is not decoded again.  The second
    transmission's LLRs of those codeblocks are replaced by noise; the server's TB results
    and bytes must equal the oracle chain, which decodes only the failed codeblocks (with
    HARQ combining) and keeps the others."""
    print("case HARQ codeblock CRC state", flush=True)
    from ocudu_hip import serve
    ctx, _ = setup
    it = serve.PUSCH_LDPC_ITERATIONS
    thr = serve.mcs_thresholds(ctx, n=64)
    trace = serve.Trace(ctx, slots=8, mean_ues=8, thresholds=thr, seed=17, snr_lo=-0.5, snr_hi=-0.2, prb_max=120)
    llr = trace.d_llr.download(np.int8, trace.d_llr.nbytes)
    models = {}
    first = {}
    for i, t in enumerate(trace.tbs):
        if t["seg"].nof_cbs >= 2:
            models[i] = _model(t)
            n = t["nsym"] * t["qm"]
            first[i] = models[i].transmit(llr[t["llr_off"][0]:t["llr_off"][0] + n], 0, True, it)
    cand = [i for i in models if 0 < first[i][1] < models[i].C]
    assert cand, "no TB with some codeblocks passing in its first transmission"
    rng = np.random.default_rng(3)
    for i in cand:
        t = trace.tbs[i]
        for c in np.flatnonzero(np.array(models[i].passed) > 0):
            noise = rng.choice(np.array([-60, 60], np.int8), t["rm"][c])
            o = t["llr_off"][1] + t["off"][c]
            llr[o:o + noise.size] = noise
            trace.d_llr.upload(noise, offset=o)
    second = {}
    for i in cand:
        t = trace.tbs[i]
        n = t["nsym"] * t["qm"]
        second[i] = models[i].transmit(llr[t["llr_off"][1]:t["llr_off"][1] + n], 2, False, it)
    assert any(w[0] == H.LH_TB_CRC_OK for w in second.values()), "no TB completed by its retransmission"
    srv = H.Server(ctx, 1 << 30, 4096, 1, it, learned=False, nof_fallback_iterations=it)
    bufs = {}
    for tx, want in ((0, first), (1, second)):
        for i in cand:
            bufs[(i, tx)] = np.zeros(trace.tbs[i]["tbs"] // 8, np.uint8)
            srv.submit(trace.request(i, tx), bufs[(i, tx)].ctypes.data)
        srv.step()
        got = {}
        while len(got) < len(cand):
            for r in srv.poll(wait=True):
                got[r.id // 2] = r
        for i in cand:
            _same(got[i], want[i], f"TB {i} transmission {tx}")
            if want[i][0] == H.LH_TB_CRC_OK:
                assert np.array_equal(bufs[(i, tx)], want[i][3])
    srv.close()


def test_invalid_request_leaves_the_server_working(setup):
    """A request that a launch could not run (N_ref shorter than the systematic part) is
    refused at submission; a valid request submitted with it is still served, equal to the
    oracle, and no state is left behind."""
    print("case invalid request", flush=True)
    from ocudu_hip import serve
    ctx, trace = setup
    it = serve.PUSCH_LDPC_ITERATIONS
    i = max(range(len(trace.tbs)), key=lambda k: trace.tbs[k]["seg"].nof_cbs)
    t = trace.tbs[i]
    srv = H.Server(ctx, 1 << 28, 1024, 1, it)
    good = trace.request(i, 0)
    bad = trace.request(i, 0)
    bad.id, bad.harq_process, bad.Nref = 99, 1000, 64
    out = np.zeros(t["tbs"] // 8, np.uint8)
    reqs = (H.TbRequest * 2)(good, bad)
    outs = (H.C.c_void_p * 2)(out.ctypes.data, None)
    with pytest.raises(H.LhError, match="N_ref"):
        srv.submit_batch(reqs, outs)
    srv.step()
    got = []
    while not got:
        got = srv.poll(wait=True)
    assert len(got) == 1 and got[0].id == good.id
    llr = trace.d_llr.download(np.int8, t["nsym"] * t["qm"], offset=t["llr_off"][0])
    w = _model(t).transmit(llr, 0, True, it)
    _same(got[0], w, "valid request")
    st = srv.stats()
    assert (st.pending, st.ready, st.slots_in_flight) == (0, 0, 0)
    srv.close()


def _harq_need(bg, Z, C):
    n = (66 if bg == 1 else 50) * Z
    return ((C * (n + ((22 if bg == 1 else 10) * Z + 7) // 8)) + 255) // 256 * 256


def test_minimum_budget_serves_the_largest_codeblock(setup):
    """The smallest budget lh_server_create accepts holds one BG1 Z=384 codeblock's HARQ
    process (soft bits and message bits, as the arena places them), and such a TB is
    served; one byte less is refused."""
    print("case minimum budget", flush=True)
    ctx, _ = setup
    need = _harq_need(1, 384, 1)
    with pytest.raises(H.LhError):
        H.Server(ctx, need - 1, 64, 1, 6)
    srv = H.Server(ctx, need, 64, 1, 6)
    m = HarqModel(1, 8192, 2, 1, 4608)
    assert (m.C, m.Z) == (1, 384)
    rng = np.random.default_rng(4)
    bits = tb_codeblocks(rng.integers(0, 256, 1024, dtype=np.uint8), m.seg)
    d, l = _tb_llrs(ctx, m, bits, 0, 0.05, 7)
    out = np.zeros(1024, np.uint8)
    r = _one(srv, _request(d, 0, 0, m, 1, 4608, 0), out)
    w = m.transmit(l, 0, True, 6)
    _same(r, w, "minimum budget")
    assert r.tb_crc_ok == H.LH_TB_CRC_OK and np.array_equal(out, w[3])
    srv.close()


def test_harq_expiry(setup):
    """rx_buffer_pool semantics (rx_buffer_pool_impl.cpp:79-125, the reference's
    rx_buffer_pool_test buffer_expire / buffer_renew_expire): a process in use has its
    expiry renewed by run_slot; once idle it holds its memory until expire_timeout_slots
    after that, then it is released, a request waiting for memory is admitted, and a
    retransmission for the expired process is rejected."""
    print("case HARQ expiry", flush=True)
    ctx, _ = setup
    seg, rm, off = H.segment_rx(1, 8192, 2, 1, 4608)
    assert (seg.nof_cbs, seg.lifting_size) == (1, 384)
    srv = H.Server(ctx, _harq_need(1, 384, 1), 64, 1, 6, expire_timeout_slots=4)  # one process fits
    d0 = H.DeviceBuffer(9216)
    d0.zero()  # all-zero LLRs: the codeblock fails and its process keeps its buffers

    def req(rid, proc, tx, slot):
        return H.TbRequest(id=rid, d_llrs=d0.ptr, harq_process=proc, new_data=int(tx == 0), base_graph=1,
                           tbs_bits=8192, modulation_order=2, nof_layers=1, nof_ch_symbols=4608, rv=2 * tx, Nref=0,
                           slot=slot)

    srv.submit(req(0, 5, 0, 0))
    srv.step()
    srv.run_slot(10)  # in use: expiry renewed to 14
    got = []
    while not got:
        got = srv.poll(wait=True)
    assert got[0].tb_crc_ok == H.LH_TB_CB_FAILED
    assert srv.stats().harq_processes == 1
    srv.submit(req(1, 6, 0, 11))  # another process: waits for memory
    srv.step()
    assert srv.poll() == [] and srv.stats().pending == 1
    srv.run_slot(13)
    assert srv.stats().pending == 1 and srv.stats().harq_processes == 1
    srv.run_slot(14)  # process 5 expires; process 6 is admitted
    assert srv.stats().pending == 0
    srv.step()
    got = []
    while not got:
        got = srv.poll(wait=True)
    assert got[0].id == 1 and got[0].tb_crc_ok == H.LH_TB_CB_FAILED
    r = _one(srv, req(2, 5, 1, 15))  # retransmission of the expired process
    assert r.tb_crc_ok == H.LH_TB_REJECTED
    srv.close()


def test_native_loop_equals_python_loop(setup):
    """The library's slot loop (lh_server_serve) against the Python loop over the same trace:
    deterministic runs give the same transmissions with the same results and bytes, and a
    real-time run gives the same result for every transmission both made."""
    print("case native slot loop", flush=True)
    from ocudu_hip import serve
    ctx, trace = setup
    py = serve.serve(ctx, trace, realtime=False, timeout_s=120)
    nat = serve.serve(ctx, trace, realtime=False, timeout_s=120, native=True)
    rt = serve.serve(ctx, trace, realtime=True, timeout_s=120, native=True)
    assert set(py[0]) == set(nat[0])
    for k, r in py[0].items():
        m = nat[0][k]
        assert (r.tb_crc_ok, r.nof_cbs_ok, r.max_iterations) == (m.tb_crc_ok, m.nof_cbs_ok, m.max_iterations), k
        if r.tb_crc_ok == H.LH_TB_CRC_OK:
            assert np.array_equal(py[3][k], nat[3][k])
    common = set(py[0]) & set(rt[0])
    assert len(common) >= len(trace.tbs)
    for k in common:
        r, m = py[0][k], rt[0][k]
        assert (r.tb_crc_ok, r.nof_cbs_ok, r.max_iterations) == (m.tb_crc_ok, m.nof_cbs_ok, m.max_iterations), k
    assert len(nat[1]) > 0 and all(x >= 0 for x in nat[1])

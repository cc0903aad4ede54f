"""GPU parity of the remaining entry points and decoder modes: standalone rate
dematcher (new data and HARQ combining, all RVs, Nref, fillers, wrap-around), device
CRC, decoder without CRC (syndrome early stop / no early stop), force decoding,
all-zero and short inputs, input lengths that are not a multiple of Z."""
import numpy as np
import pytest

import ocudu_hip as H
import oracle as O
from gpu_util import ROW_MODES, decode_batch_gpu, make_ctx, require_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=ROW_MODES)
def ctx(request):
    c = make_ctx(4096, request.param)
    yield c
    c.close()


def test_rate_dematch_sweep(ctx):
    print("case dematch sweep", flush=True)
    rng = np.random.default_rng(0)
    descs, ins, olds, exp = [], [], [], {True: [], False: []}
    cw_off = rm_off = 0
    cases = []
    for bg in (1, 2):
        for Z in (2, 11, 36, 104, 384):
            for rv in range(4):
                for Qm in (1, 2, 4, 6, 8):
                    nsys = (O.K_b(bg) - 2) * Z
                    F = int(rng.choice([0, 2, 4, 8])) if nsys > 8 else 0
                    N = O.N_short(bg) * Z
                    Nref = int(rng.choice([0, int(N * 0.75)]))
                    if Nref and Nref <= nsys:
                        Nref = 0
                    E = max(Qm, int(N * rng.uniform(0.2, 2.5)) // Qm * Qm)
                    cases.append((bg, Z, rv, Qm, F, Nref, E))
    cw_parts = []
    for (bg, Z, rv, Qm, F, Nref, E) in cases:
        N = O.N_short(bg) * Z
        x = rng.integers(-120, 121, E).astype(np.int8)
        old = rng.integers(-120, 121, N).astype(np.int8)
        descs.append(H.RmDesc(cw_off, rm_off, E, bg, Z, rv, Qm, Nref, F))
        cw_parts.append((cw_off, x))
        olds.append((rm_off, old))
        exp[True].append(O.rate_dematch(x, bg, Z, rv=rv, Qm=Qm, Nref=Nref, F=F, new_data=True, harq=old))
        exp[False].append(O.rate_dematch(x, bg, Z, rv=rv, Qm=Qm, Nref=Nref, F=F, new_data=False, harq=old))
        cw_off += E
        rm_off += N
    cw = np.zeros(cw_off, np.int8)
    for o, a in cw_parts:
        cw[o:o + a.size] = a
    oldbuf = np.zeros(rm_off, np.int8)
    for o, a in olds:
        oldbuf[o:o + a.size] = a
    d_cw = H.DeviceBuffer.from_array(cw)
    for new_data in (True, False):
        d_rm = H.DeviceBuffer.from_array(oldbuf)
        ctx.rate_dematch(descs, d_cw, d_rm, new_data=new_data)
        got = d_rm.download(np.int8, rm_off)
        for k, (o, old) in enumerate(olds):
            assert np.array_equal(got[o:o + old.size], exp[new_data][k]), f"case {cases[k]} new_data={new_data}"


def test_crc_calculate(ctx):
    print("case crc", flush=True)
    rng = np.random.default_rng(1)
    sizes = [1, 7, 8, 9, 24, 100, 1000, 3824, 8424, 8448]
    bufs = [rng.integers(0, 256, (s + 7) // 8).astype(np.uint8) for s in sizes]
    offs = np.cumsum([0] + [b.size for b in bufs])
    allb = np.concatenate(bufs)
    d_b = H.DeviceBuffer.from_array(allb)
    d_o = H.DeviceBuffer(4 * len(sizes))
    for poly in (O.CRC24A, O.CRC24B, O.CRC16):
        ctx.crc_calculate(poly, [int(x) for x in offs[:-1]], sizes, d_b, d_o)
        got = d_o.download(np.uint32, len(sizes))
        for k, s in enumerate(sizes):
            assert got[k] == O.crc(poly, bufs[k], s), (poly, s)


def _llrs(bg, Z, snr, seed, Qm=2, frac=1.0):
    K = O.K_b(bg) * Z
    poly, cb = (O.CRC24B, 24) if K > 48 else (O.CRC16, 16)
    N = O.N_short(bg) * Z
    E = max(Qm, int(N * frac) // Qm * Qm)
    _, le = O.make_codeblocks(1, bg, Z, E, Qm, snr, seed=seed, crc_poly=poly, crc_bits=cb)
    return O.rate_dematch(le[0], bg, Z, Qm=Qm), poly, cb


@pytest.mark.parametrize("flags", [0, H.LH_EARLY_STOP_SYNDROME])
def test_no_crc_modes(ctx, flags):
    print(f"case no-crc flags={flags}", flush=True)
    llrs, cfgs = [], []
    for k, (bg, Z, snr) in enumerate([(1, 384, 2.0), (1, 384, 1.0), (2, 208, 2.5), (1, 52, 3.0), (2, 13, 1.5),
                                      (2, 384, 4.0), (1, 7, 2.0), (2, 64, 0.5)]):
        a, _, cb = _llrs(bg, Z, snr, 50 + k, frac=0.9)
        llrs.append(a)
        cfgs.append((bg, Z, 0, cb, 5))
    res, bits = decode_batch_gpu(ctx, llrs, cfgs, [0] * len(llrs), flags)
    for i, a in enumerate(llrs):
        bg, Z, F, cb, mi = cfgs[i]
        r_o, b_o = O.decode(a, bg, Z, F=F, crc_poly=0, max_iter=mi, crc_bits=cb,
                            es_syn=bool(flags & H.LH_EARLY_STOP_SYNDROME))
        assert res[i] == r_o and np.array_equal(bits[i], b_o), f"cb {i}"


def test_special_inputs(ctx):
    print("case special inputs", flush=True)
    llrs, cfgs, polys, force = [], [], [], []
    # all zero input -> no decode; short input; input only up to K*Z - 1
    for bg, Z in ((1, 384), (2, 10)):
        N = O.N_short(bg) * Z
        K = O.K_b(bg) * Z
        llrs.append(np.zeros(N, np.int8))
        cfgs.append((bg, Z, 0, 24, 4))
        polys.append(O.CRC24B)
        a, p, cb = _llrs(bg, Z, 3.0, 7)
        b = a.copy()
        b[K - 2 * Z - 1:] = 0  # last nonzero before the message length
        llrs.append(b)
        cfgs.append((bg, Z, 0, cb, 4))
        polys.append(p)
        llrs.append(b.copy())
        cfgs.append((bg, Z, 0, cb, 4))
        polys.append(0)  # no CRC: output must be all ones
    res, bits = decode_batch_gpu(ctx, llrs, cfgs, polys)
    for i, a in enumerate(llrs):
        bg, Z, F, cb, mi = cfgs[i]
        out0 = np.zeros(H.msg_bytes(bg, Z), np.uint8)
        r_o, b_o = O.decode(a, bg, Z, F=F, crc_poly=polys[i], max_iter=mi, crc_bits=cb, out=out0)
        assert res[i] == r_o and np.array_equal(bits[i], b_o), f"cb {i}"
    # force decoding of the short inputs
    res, bits = decode_batch_gpu(ctx, llrs, cfgs, polys, H.LH_FORCE_DECODING)
    for i, a in enumerate(llrs):
        bg, Z, F, cb, mi = cfgs[i]
        r_o, b_o = O.decode(a, bg, Z, F=F, crc_poly=polys[i], max_iter=mi, crc_bits=cb, force=True,
                            out=np.zeros(H.msg_bytes(bg, Z), np.uint8))
        assert res[i] == r_o and np.array_equal(bits[i], b_o), f"forced cb {i}"


def test_ragged_input_lengths(ctx):
    """Input lengths between (K_b+2)Z and N_short*Z that are not multiples of Z: the last
    partial node is loaded unclamped (the reference's tail rule)."""
    print("case ragged", flush=True)
    rng = np.random.default_rng(9)
    llrs, cfgs, polys = [], [], []
    for k in range(16):
        bg = 1 + k % 2
        Z = int(rng.choice([5, 24, 56, 144, 384]))
        a, p, cb = _llrs(bg, Z, 4.0, 300 + k)
        lo, hi = (O.K_b(bg) + 2) * Z, O.N_short(bg) * Z
        n = int(rng.integers(lo, hi))
        b = a[:n].copy()
        b[-(n % Z or 1):] = rng.choice([-127, 127, -100, 90], size=(n % Z or 1))
        llrs.append(b)
        cfgs.append((bg, Z, 0, cb, 6))
        polys.append(p)
    res, bits = decode_batch_gpu(ctx, llrs, cfgs, polys)
    for i, a in enumerate(llrs):
        bg, Z, F, cb, mi = cfgs[i]
        r_o, b_o = O.decode(a, bg, Z, F=F, crc_poly=polys[i], max_iter=mi, crc_bits=cb)
        assert res[i] == r_o and np.array_equal(bits[i], b_o), f"cb {i} n={a.size} Z={Z}"


def test_fillers_and_harq_combining(ctx):
    """PUSCH batch with fillers, rv0 then rv2 retransmission combined in the HARQ buffer."""
    print("case harq", flush=True)
    rng = np.random.default_rng(4)
    n = 12
    specs = []
    for i in range(n):
        bg = 1 + i % 2
        Z = int(rng.choice([32, 96, 208, 384]))
        F = int(rng.choice([0, 16, 40]))
        Qm = int(rng.choice([2, 4, 6]))
        N = O.N_short(bg) * Z
        E = int(N * 0.45) // Qm * Qm
        specs.append((bg, Z, F, Qm, E, N))
    rm_tot = sum(s[5] for s in specs)
    d_rm = H.DeviceBuffer(rm_tot)
    bits_tot = sum(H.msg_bytes(s[0], s[1]) for s in specs)
    d_bits = H.DeviceBuffer(bits_tot)
    d_bits.zero()
    d_res = H.DeviceBuffer(4 * n)
    harq_exp = [None] * n
    for rv, new_data in ((0, True), (2, False)):
        cbs, parts = [], []
        cw_off = rm_off = bo = 0
        expect = []
        for i, (bg, Z, F, Qm, E, N) in enumerate(specs):
            _, le = O.make_codeblocks(1, bg, Z, E, Qm, -1.0, seed=700 + i, rv=rv, F=F)
            cbs.append(H.PuschCbDesc(H.RmDesc(cw_off, rm_off, E, bg, Z, rv, Qm, 0, F), bo, 24, O.CRC24B))
            parts.append((cw_off, le[0]))
            harq_exp[i] = O.rate_dematch(le[0], bg, Z, rv=rv, Qm=Qm, F=F, new_data=new_data,
                                         harq=harq_exp[i] if harq_exp[i] is not None else np.zeros(N, np.int8))
            r_o, b_o = O.decode(harq_exp[i], bg, Z, F=F, crc_poly=O.CRC24B, max_iter=6)
            expect.append((rm_off, N, bo, H.msg_bytes(bg, Z), r_o, b_o))
            cw_off += E
            rm_off += N
            bo += H.msg_bytes(bg, Z)
        cw = np.zeros(cw_off, np.int8)
        for o, a in parts:
            cw[o:o + a.size] = a
        d_cw = H.DeviceBuffer.from_array(cw)
        ctx.pusch_decode_batch(cbs, d_cw, d_rm, d_bits, d_res, new_data=new_data, nof_ldpc_iterations=6)
        res = d_res.download(np.int32, n)
        rm = d_rm.download(np.int8, rm_tot)
        bits = d_bits.download(np.uint8, bits_tot)
        for i, (ro, N, b0, nb, r_o, b_o) in enumerate(expect):
            assert np.array_equal(rm[ro:ro + N], harq_exp[i]), f"rv{rv} cb {i} HARQ"
            assert res[i] == r_o and np.array_equal(bits[b0:b0 + nb], b_o), f"rv{rv} cb {i}"


def test_pusch_without_early_stop(ctx):
    print("case no early stop", flush=True)
    llr_list, cbs, parts, expect = [], [], [], []
    cw_off = rm_off = bo = 0
    for i, (bg, Z, Qm, snr) in enumerate([(1, 384, 2, 2.0), (2, 120, 4, 6.0), (1, 40, 6, 9.0), (2, 384, 2, 1.0)]):
        N = O.N_short(bg) * Z
        E = int(N * 0.6) // Qm * Qm
        _, le = O.make_codeblocks(1, bg, Z, E, Qm, snr, seed=900 + i)
        cbs.append(H.PuschCbDesc(H.RmDesc(cw_off, rm_off, E, bg, Z, 0, Qm, 0, 0), bo, 24, O.CRC24B))
        parts.append((cw_off, le[0]))
        rm = O.rate_dematch(le[0], bg, Z, Qm=Qm)
        # pusch_codeblock_decoder.cpp:47-57: decode without CRC, then CRC on success
        r_o, b_o = O.decode(rm, bg, Z, crc_poly=0, max_iter=5)
        if r_o and O.crc(O.CRC24B, b_o, O.K_b(bg) * Z) != 0:
            r_o = 0
        expect.append((bo, H.msg_bytes(bg, Z), r_o, b_o))
        cw_off += E
        rm_off += N
        bo += H.msg_bytes(bg, Z)
    cw = np.zeros(cw_off, np.int8)
    for o, a in parts:
        cw[o:o + a.size] = a
    d_cw = H.DeviceBuffer.from_array(cw)
    d_rm, d_bits, d_res = H.DeviceBuffer(rm_off), H.DeviceBuffer(bo), H.DeviceBuffer(16)
    d_bits.zero()
    ctx.pusch_decode_batch(cbs, d_cw, d_rm, d_bits, d_res, new_data=True, use_early_stop=False,
                           nof_ldpc_iterations=5)
    res = d_res.download(np.int32, 4)
    bits = d_bits.download(np.uint8, bo)
    for i, (b0, nb, r_o, b_o) in enumerate(expect):
        assert res[i] == r_o and np.array_equal(bits[b0:b0 + nb], b_o), f"cb {i}"

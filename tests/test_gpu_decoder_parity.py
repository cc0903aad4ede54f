"""Bit-exact parity of the HIP decoder with the oracle (restatement of the reference
generic decoder): same hard bits, same iteration count, same CRC verdict."""
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


def _check(ctx, llrs, cfgs, polys, flags=0, force=False, es_syn=False):
    res, bits = decode_batch_gpu(ctx, llrs, cfgs, polys, flags)
    for i, (a, c) in enumerate(zip(llrs, cfgs)):
        bg, Z, F, cb, mi = c
        r_o, b_o = O.decode(a, bg, Z, F=F, crc_poly=polys[i], max_iter=mi, crc_bits=cb, force=force,
                            es_syn=es_syn, out=np.zeros(H.msg_bytes(bg, Z), np.uint8))
        assert res[i] == r_o, f"cb {i} (BG{bg} Z{Z}): iterations {res[i]} vs oracle {r_o}"
        assert np.array_equal(bits[i], b_o), f"cb {i} (BG{bg} Z{Z}): hard bits differ"
    return res


def test_cfg1_bg1_z384_qpsk_2db(ctx):
    # 64 codeblocks BG1 Z=384, E=16896, QPSK, Es/N0 = 2 dB, 6 iterations, CRC24B.
    print("case cfg1", flush=True)
    msgs, llr_e = O.make_codeblocks(64, 1, 384, 16896, 2, 2.0, seed=1)
    rm = [O.rate_dematch(llr_e[c], 1, 384, Qm=2) for c in range(64)]
    res = _check(ctx, rm, [(1, 384, 0, 24, 6)] * 64, [O.CRC24B] * 64)
    assert (res > 0).sum() >= 1


@pytest.mark.parametrize("bg", [1, 2])
def test_all_lifting_sizes(ctx, bg):
    print(f"case all Z bg{bg}", flush=True)
    llrs, cfgs, polys = [], [], []
    for k, Z in enumerate(O.LIFTING_SIZES):
        K = O.K_b(bg) * Z
        E = ((O.N_short(bg) * Z * 2 // 3) // 4) * 4
        poly, cb = (O.CRC24B, 24) if K > 48 else (O.CRC16, 16)
        for snr, seed in ((0.5, 3 * k), (3.0, 3 * k + 1)):
            _, le = O.make_codeblocks(1, bg, Z, E, 4, snr, seed=seed + 1000 * bg, crc_poly=poly, crc_bits=cb)
            llrs.append(O.rate_dematch(le[0], bg, Z, Qm=4))
            cfgs.append((bg, Z, 0, cb, 8))
            polys.append(poly)
    _check(ctx, llrs, cfgs, polys)


def test_pusch_batch_fused_dematch(ctx):
    print("case pusch batch", flush=True)
    rng = np.random.default_rng(5)
    cbs, cw_parts, expect = [], [], []
    cw_off = rm_off = bits_off = 0
    for i in range(24):
        bg = 1 + (i % 2)
        Z = int(rng.choice(O.LIFTING_SIZES[20:]))
        Qm = int(rng.choice([2, 4, 6, 8]))
        F = int(rng.integers(0, 3)) * 8 if Z >= 16 else 0
        N = O.N_short(bg) * Z
        E = int(N * rng.uniform(0.4, 1.3)) // Qm * Qm
        rv = int(rng.integers(0, 4))
        _, le = O.make_codeblocks(1, bg, Z, E, Qm, 6.0, seed=100 + i, rv=rv, F=F)
        cbs.append(H.PuschCbDesc(H.RmDesc(cw_off, rm_off, E, bg, Z, rv, Qm, 0, F), bits_off, 24, O.CRC24B))
        cw_parts.append((cw_off, le[0]))
        rm_exp = O.rate_dematch(le[0], bg, Z, rv=rv, Qm=Qm, F=F)
        r_o, b_o = O.decode(rm_exp, bg, Z, F=F, crc_poly=O.CRC24B, max_iter=6)
        expect.append((rm_off, N, rm_exp, bits_off, H.msg_bytes(bg, Z), r_o, b_o))
        cw_off += E
        rm_off += N
        bits_off += H.msg_bytes(bg, Z)
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
    res = d_res.download(np.int32, len(cbs))
    rm = d_rm.download(np.int8, rm_off)
    bits = d_bits.download(np.uint8, bits_off)
    for i, (ro, N, rm_exp, bo, nb, r_o, b_o) in enumerate(expect):
        assert np.array_equal(rm[ro:ro + N], rm_exp), f"cb {i}: HARQ buffer differs"
        assert res[i] == r_o, f"cb {i}: iterations {res[i]} vs {r_o}"
        assert np.array_equal(bits[bo:bo + nb], b_o), f"cb {i}: bits differ"

"""The headline workload on the device against the oracle: BG1 Z=384, 256-QAM, E=9088
(code rate ~0.93, the four-layer (K_b + 4) Z floor of ldpc_decoder_impl.cpp:90-99), rv0
new data, 8 iterations, CRC24B early stop, LLRs made by the device generator near the
decoding threshold (a mix of passing and failing codeblocks).  Every codeblock is checked
bit-exactly: the dematched HARQ buffer, the hard bits (also of failing codeblocks), the
iteration count and the CRC verdict.  One full 152-codeblock slot in both row modes, and
one 2432-codeblock launch (16 slots, the benchmark's step) in the row-pair mode."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import ocudu_hip as H
import oracle as O
from gpu_util import require_gpu

pytestmark = pytest.mark.gpu

BG, Z, E, QM, ITERS = 1, 384, 9088, 8, 8
N = 66 * Z
MSG_B = 22 * Z // 8


def _run(n_cb, snrs, rows, seed):
    """Decodes n_cb codeblocks (Es/N0 cycling over snrs) with the row mode forced; returns
    the LLRs and the device's HARQ buffers, bits and results."""
    require_gpu()
    ctx = H.Context(0, n_cb)
    ctx.tune_rows(BG, Z, rows)
    seeds = np.random.default_rng(seed).integers(0, 2 ** 63, (n_cb, 2))
    d_cw, d_msg = H.DeviceBuffer(n_cb * E), H.DeviceBuffer(n_cb * MSG_B)
    gen = [H.GenDesc(i * E, i * MSG_B, int(seeds[i, 0]), int(seeds[i, 1]), float(10.0 ** (-snrs[i % len(snrs)] / 10)),
                     E, BG, Z, 0, QM, 0, 0, H.LH_CRC24B) for i in range(n_cb)]
    ctx.gen_codeblocks(gen, d_cw, d_msg)
    d_rm, d_bits, d_res = H.DeviceBuffer(n_cb * N), H.DeviceBuffer(n_cb * MSG_B), H.DeviceBuffer(4 * n_cb)
    d_bits.zero()
    cbs = [H.PuschCbDesc(H.RmDesc(i * E, i * N, E, BG, Z, 0, QM, 0, 0), i * MSG_B, 24, H.LH_CRC24B)
           for i in range(n_cb)]
    ctx.pusch_decode_batch(cbs, d_cw, d_rm, d_bits, d_res, nof_ldpc_iterations=ITERS)
    out = (d_cw.download(np.int8, n_cb * E).reshape(n_cb, E), d_rm.download(np.int8, n_cb * N).reshape(n_cb, N),
           d_bits.download(np.uint8, n_cb * MSG_B).reshape(n_cb, MSG_B), d_res.download(np.int32, n_cb),
           d_msg.download(np.uint8, n_cb * MSG_B).reshape(n_cb, MSG_B))
    ctx.close()
    return out


def _oracle(llr_e):
    rm = O.rate_dematch(llr_e, BG, Z, Qm=QM)
    r, b = O.decode(rm, BG, Z, crc_poly=O.CRC24B, max_iter=ITERS)
    return rm, r, b


def _check(llrs, rm, bits, res, msgs):
    with ThreadPoolExecutor(8) as ex:  # the oracle's C calls release the GIL
        want = list(ex.map(_oracle, llrs))
    for i, (rm_o, r_o, b_o) in enumerate(want):
        assert np.array_equal(rm[i], rm_o), f"cb {i}: HARQ buffer differs"
        assert np.flatnonzero(rm_o).max() < 24 * Z  # four layers
        assert res[i] == r_o, f"cb {i}: iterations {res[i]} vs oracle {r_o}"
        assert np.array_equal(bits[i], b_o), f"cb {i}: hard bits differ (result {r_o})"
        if r_o:
            assert np.array_equal(bits[i], msgs[i]), f"cb {i}: passing codeblock decoded to another message"
    return np.array([w[1] for w in want])


@pytest.mark.parametrize("rows", [0, 1])
def test_slot_152(rows):
    print(f"case cfg2 slot rows={rows}", flush=True)
    llrs, rm, bits, res, msgs = _run(152, (24.9, 25.1, 25.3, 25.5), rows, seed=7)
    r = _check(llrs, rm, bits, res, msgs)
    assert 10 <= (r == 0).sum() <= 142, "want both passing and failing codeblocks"
    assert (r == ITERS).sum() >= 1


def test_step_2432_row_pairs():
    print("case cfg2 step 2432 row pairs", flush=True)
    llrs, rm, bits, res, msgs = _run(2432, (25.2, 25.35, 25.5, 26.0), 1, seed=8)
    r = _check(llrs, rm, bits, res, msgs)
    assert (r == 0).sum() >= 50 and (r > 0).sum() >= 1000
